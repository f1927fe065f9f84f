// Transfer-matrix read-outs of device-resident MPS chains on gfx950: left environments,
// one- and two-site expectation values and the overlap of two chains.  These have no
// counterpart in the reference, whose only read-outs are the dense statevector and the
// host-side sampling (mps_core.py:354-461); they replace the host copy of the chain that
// such observables would otherwise need.
//
//  * One transfer step maps an environment E (x_l x y_l) through site X of one chain
//    (bra) and site Y of another (ket), with an optional one-site operator O between them:
//        E'[a', b'] = sum_{s,t} O[s][t] sum_{a,b} conj(X[a, s, a']) E[a, b] Y[b, t, b'].
//    Workgroup (ta, tb) computes the 16 x 16 tile a' in [16 ta, 16 ta + 16), b' in
//    [16 tb, 16 tb + 16): first U_s[a, b'] = sum_t O[s][t] sum_b E[a, b] Y[b, t, b'] into
//    LDS for every a, then the tile from the tile's columns of X, also staged in LDS.
//    grid.z runs independent steps (one per site) side by side.
//  * E is divided by a scale taken from the input environment: its trace (environments of
//    one chain, Hermitian and positive) or the sum of |Re| + |Im| of its entries (two
//    chains); workgroup (0, 0) records log(scale).  Sums run in index order and are folded
//    in a fixed tree, so results are deterministic.
//  * Left environments L_0 = [[1]], L_{i+1} = transfer(A_i, L_i, A_i), each stored divided
//    by its own trace, mirror the right environments of the sampler (mps_sample.hip).
//  * <O_i> = close(transfer(A_i, L_i, A_i; O), R_{i+1}) / close(transfer(A_i, L_i, A_i),
//    R_{i+1}) with close(M, R) = sum_{a', b'} M[a', b'] R[b', a'], for all sites at once;
//    pairs <O_i O'_{i+1}> take a second transfer through site i+1 and close with R_{i+2}.
#include <algorithm>
#include <cmath>
#include <vector>

#include "qsim_internal.h"

namespace qs {

namespace {

constexpr int kTile = 16;
constexpr int kMaxChi = 128;
constexpr int kThreads = kTile * kTile;
constexpr int kExpectChunk = 128;     // sites per chunk of the expectation read-outs

__device__ __forceinline__ cd2 cmadd(cd2 acc, cd2 a, cd2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}
__device__ __forceinline__ cd2 cjmadd(cd2 acc, cd2 a, cd2 b) {   // acc + conj(a) * b
    acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(-a.y, b.x, acc.y));
    return acc;
}

// Block-wide sum in a fixed order: thread partials folded by a tree in LDS.
__device__ __forceinline__ double block_sum(double v, double* red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

struct Op2 {
    cd2 m[2][2];
};

// One transfer step per grid.z item z (site = site0 + z).  The input environment of
// item z is ein + z * ein_step, its output eout + z * eout_step; both have row stride ld.
// scale_mode 0: none, 1: trace of the input, 2: sum of |Re| + |Im| of the input.
__global__ void __launch_bounds__(kThreads)
mps_transfer_kernel(const cd2* __restrict__ xs, const int32_t* __restrict__ xbond,
                    int64_t xstride, const cd2* __restrict__ ys,
                    const int32_t* __restrict__ ybond, int64_t ystride, int site0,
                    const cd2* __restrict__ ein, int64_t ein_step, cd2* __restrict__ eout,
                    int64_t eout_step, int ld, Op2 op, int use_op, int scale_mode,
                    double* __restrict__ logs) {
    __shared__ cd2 U[2][kMaxChi][kTile];
    __shared__ cd2 Xt[2][kMaxChi][kTile];
    __shared__ double red[kThreads];
    const int z = blockIdx.z;
    const int site = site0 + z;
    const int xl = xbond[site], xr = xbond[site + 1];
    const int yl = ybond[site], yr = ybond[site + 1];
    const int a0 = blockIdx.x * kTile, b0 = blockIdx.y * kTile;
    if (a0 >= xr || b0 >= yr) return;                  // uniform per workgroup
    const cd2* X = xs + (size_t)site * (size_t)xstride;   // (xl, 2, xr)
    const cd2* Y = ys + (size_t)site * (size_t)ystride;   // (yl, 2, yr)
    const cd2* E = ein + (size_t)z * (size_t)ein_step;    // (xl x yl)
    cd2* O = eout + (size_t)z * (size_t)eout_step;        // (xr x yr)
    const int tid = threadIdx.x;
    double inv = 1.0;
    if (scale_mode != 0) {
        double part = 0.0;
        if (scale_mode == 1) {
            for (int k = tid; k < xl; k += kThreads) part += E[(size_t)k * ld + k].x;
        } else {
            for (int k = tid; k < xl * yl; k += kThreads) {
                const cd2 v = E[(size_t)(k / yl) * ld + (k % yl)];
                part += fabs(v.x) + fabs(v.y);
            }
        }
        const double sc = block_sum(part, red);
        const bool ok = sc > 1e-300 && isfinite(sc);
        inv = ok ? 1.0 / sc : 1.0;
        if (logs != nullptr && blockIdx.x == 0 && blockIdx.y == 0 && tid == 0)
            logs[site] = ok ? log(sc) : 0.0;
    }
    // U_s[a, b'] for the tile's b' columns, and X[a, s, a'] for the tile's a' columns
    for (int idx = tid; idx < xl * kTile; idx += kThreads) {
        const int a = idx / kTile, bl = idx - a * kTile;
        const int bp = b0 + bl, ap = a0 + bl;
        cd2 v0 = mk2(0.0, 0.0), v1 = mk2(0.0, 0.0);
        if (bp < yr) {
            const cd2* erow = E + (size_t)a * ld;
            for (int b = 0; b < yl; ++b) {
                const cd2 e = erow[b];
                v0 = cmadd(v0, e, Y[((size_t)b * 2 + 0) * yr + bp]);
                v1 = cmadd(v1, e, Y[((size_t)b * 2 + 1) * yr + bp]);
            }
        }
        v0 = mk2(v0.x * inv, v0.y * inv);
        v1 = mk2(v1.x * inv, v1.y * inv);
        if (use_op) {
            cd2 u0 = mk2(0.0, 0.0), u1 = mk2(0.0, 0.0);
            u0 = cmadd(cmadd(u0, op.m[0][0], v0), op.m[0][1], v1);
            u1 = cmadd(cmadd(u1, op.m[1][0], v0), op.m[1][1], v1);
            U[0][a][bl] = u0;
            U[1][a][bl] = u1;
        } else {
            U[0][a][bl] = v0;
            U[1][a][bl] = v1;
        }
        Xt[0][a][bl] = ap < xr ? X[((size_t)a * 2 + 0) * xr + ap] : mk2(0.0, 0.0);
        Xt[1][a][bl] = ap < xr ? X[((size_t)a * 2 + 1) * xr + ap] : mk2(0.0, 0.0);
    }
    __syncthreads();
    const int al = tid / kTile, bl = tid % kTile;
    const int ap = a0 + al, bp = b0 + bl;
    if (ap < xr && bp < yr) {
        cd2 acc = mk2(0.0, 0.0);
        for (int s = 0; s < 2; ++s)
            for (int a = 0; a < xl; ++a) acc = cjmadd(acc, Xt[s][a][al], U[s][a][bl]);
        O[(size_t)ap * ld + bp] = acc;
    }
}

// E_0 = [[1]] (one environment at `e`).
__global__ void mps_unit_env_kernel(cd2* __restrict__ e) {
    if (threadIdx.x == 0 && blockIdx.x == 0) e[0] = mk2(1.0, 0.0);
}

// Environment i of `envs` (dimension dims[i] from the bond array, offset `first`) divided
// by its own trace; one workgroup per environment.
__global__ void __launch_bounds__(kThreads)
mps_trace_normalise_kernel(const int32_t* __restrict__ bond, int first, int count, int ld,
                           cd2* __restrict__ envs) {
    __shared__ double red[kThreads];
    const int i = first + blockIdx.x;
    if ((int)blockIdx.x >= count) return;
    const int dim = bond[i];
    cd2* R = envs + (size_t)i * ld * ld;
    double part = 0.0;
    for (int k = threadIdx.x; k < dim; k += kThreads) part += R[(size_t)k * ld + k].x;
    const double tr = block_sum(part, red);
    const double s = tr > 1e-300 ? 1.0 / tr : 1.0;
    for (int idx = threadIdx.x; idx < dim * dim; idx += kThreads) {
        const int r = idx / dim, c = idx - r * dim;
        const cd2 v = R[(size_t)r * ld + c];
        R[(size_t)r * ld + c] = mk2(v.x * s, v.y * s);
    }
}

// out[z] = sum_{a', b'} M_z[a', b'] R_z[b', a'] with M_z (dims bond[site+span]^2 of the
// chain) at m + z * m_step and R_z the right environment of site site0 + z + span.
__global__ void __launch_bounds__(kThreads)
mps_close_kernel(const int32_t* __restrict__ bond, int site0, int span,
                 const cd2* __restrict__ m, int64_t m_step, const cd2* __restrict__ renvs,
                 int ld, cd2* __restrict__ out) {
    __shared__ double red[kThreads];
    const int z = blockIdx.x;
    const int k = site0 + z + span;
    const int dim = bond[k];
    const cd2* M = m + (size_t)z * (size_t)m_step;
    const cd2* R = renvs + (size_t)k * ld * ld;
    double re = 0.0, im = 0.0;
    for (int idx = threadIdx.x; idx < dim * dim; idx += kThreads) {
        const int a = idx / dim, b = idx - a * dim;
        const cd2 x = M[(size_t)a * ld + b], y = R[(size_t)b * ld + a];
        re = fma(x.x, y.x, fma(-x.y, y.y, re));
        im = fma(x.x, y.y, fma(x.y, y.x, im));
    }
    re = block_sum(re, red);
    im = block_sum(im, red);
    if (threadIdx.x == 0) out[z] = mk2(re, im);
}

// out[i] = Re(num[i]) / Re(den[i]) for i < count.
__global__ void mps_ratio_kernel(const cd2* __restrict__ num, const cd2* __restrict__ den,
                                 int count, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = den[i].x != 0.0 ? num[i].x / den[i].x : 0.0;
}

// Eigenvalues of the Hermitian (dim x dim) matrix H (row stride ld, in global memory,
// overwritten), dim <= kMaxChi, by cyclic two-sided Jacobi: each round rotates dim/2
// disjoint index pairs of a round-robin tournament (a dummy index pads odd dims).  Pair
// (p, q) with H[p][q] = |h| e^{i phi} uses J = diag(1, e^{-i phi}) R, R the real rotation
// [[c, s], [-s, c]] that diagonalises [[a, |h|], [|h|, b]] (t = sign(tau) / (|tau| +
// sqrt(1 + tau^2)), tau = (b - a) / (2|h|)); disjoint rotations commute, so a round
// applies every column rotation and then every row rotation.  A pair is skipped when
// |h| <= 1e-15 sqrt(|a b|) or |h| <= 1e-300; sweeps end when a sweep skips every pair.
__device__ void wg_hermitian_eigenvalues(cd2* H, int dim, int ld, double* evals,
                                         double* sh_c, double* sh_s, cd2* sh_e, int* sh_p,
                                         int* sh_q, int* sh_any) {
    const int tid = threadIdx.x;
    const int np = (dim + 1) & ~1;                    // players incl. the dummy
    const int half = np / 2;
    for (int sweep = 0; sweep < 60; ++sweep) {
        if (tid == 0) *sh_any = 0;
        __syncthreads();
        for (int r = 0; r < np - 1; ++r) {
            if (tid < half) {
                // round-robin: player 0 fixed, the others rotate by r
                auto player = [&](int k) { return k == 0 ? 0 : 1 + (k - 1 + r) % (np - 1); };
                int p = player(tid), q = player(np - 1 - tid);
                if (p > q) { const int x = p; p = q; q = x; }
                double c = 1.0, sn = 0.0;
                cd2 e = mk2(1.0, 0.0);
                if (q < dim) {
                    const double a = H[(size_t)p * ld + p].x, b = H[(size_t)q * ld + q].x;
                    const cd2 h = H[(size_t)p * ld + q];
                    const double g = hypot(h.x, h.y);
                    if (g > 1e-300 && g > 1e-15 * sqrt(fabs(a * b))) {
                        const double tau = (b - a) / (2.0 * g);
                        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        sn = t * c;
                        e = mk2(h.x / g, -h.y / g);          // e^{-i phi}
                        *sh_any = 1;
                    }
                }
                sh_c[tid] = c;
                sh_s[tid] = sn;
                sh_e[tid] = e;
                sh_p[tid] = p;
                sh_q[tid] = q;
            }
            __syncthreads();
            // columns: col p <- c col p - s e col q,  col q <- s col p + c e col q
            for (int idx = tid; idx < dim * half; idx += blockDim.x) {
                const int row = idx / half, k = idx - row * half;
                const int p = sh_p[k], q = sh_q[k];
                if (q >= dim || sh_s[k] == 0.0) continue;
                const double c = sh_c[k], sn = sh_s[k];
                const cd2 e = sh_e[k];
                const cd2 x = H[(size_t)row * ld + p], y = H[(size_t)row * ld + q];
                const cd2 ey = mk2(e.x * y.x - e.y * y.y, e.x * y.y + e.y * y.x);
                H[(size_t)row * ld + p] = mk2(c * x.x - sn * ey.x, c * x.y - sn * ey.y);
                H[(size_t)row * ld + q] = mk2(sn * x.x + c * ey.x, sn * x.y + c * ey.y);
            }
            __syncthreads();
            // rows: row p <- c row p - s conj(e) row q,  row q <- s row p + c conj(e) row q
            for (int idx = tid; idx < dim * half; idx += blockDim.x) {
                const int col = idx / half, k = idx - col * half;
                const int p = sh_p[k], q = sh_q[k];
                if (q >= dim || sh_s[k] == 0.0) continue;
                const double c = sh_c[k], sn = sh_s[k];
                const cd2 ec = mk2(sh_e[k].x, -sh_e[k].y);
                const cd2 x = H[(size_t)p * ld + col], y = H[(size_t)q * ld + col];
                const cd2 ey = mk2(ec.x * y.x - ec.y * y.y, ec.x * y.y + ec.y * y.x);
                H[(size_t)p * ld + col] = mk2(c * x.x - sn * ey.x, c * x.y - sn * ey.y);
                H[(size_t)q * ld + col] = mk2(sn * x.x + c * ey.x, sn * x.y + c * ey.y);
            }
            __syncthreads();
        }
        const int any = *sh_any;
        __syncthreads();
        if (!any) break;
    }
    for (int k = tid; k < dim; k += blockDim.x) evals[k] = H[(size_t)k * ld + k].x;
    __syncthreads();
}

// Entanglement entropy of bond b = bonds[blockIdx.x] by the reference's definition
// (mps_core.py:465-483): the singular values S of the sites 0..b contracted into one
// (2^(b+1) x chi) matrix M, S > 1e-15 kept, p = S^2 / sum S^2, -sum p log p.  S^2 are
// the eigenvalues of the Gram matrix M^H M = tr(G) L_{b+1}, with L_{b+1} the
// trace-normalised left environment and log tr(G) = sum_{j <= b+1} log_traces[j] (the
// traces the environment recursion divided out).  The environment is copied into the
// workspace slot of the workgroup and diagonalised there.
__global__ void __launch_bounds__(kThreads)
mps_entropy_kernel(const int32_t* __restrict__ bond, const cd2* __restrict__ lenvs, int ld,
                   const double* __restrict__ log_traces, const int32_t* __restrict__ bonds,
                   cd2* __restrict__ work, double* __restrict__ evals_out,
                   double* __restrict__ entropy) {
    __shared__ double sh_c[kMaxChi / 2 + 1], sh_s[kMaxChi / 2 + 1];
    __shared__ cd2 sh_e[kMaxChi / 2 + 1];
    __shared__ int sh_p[kMaxChi / 2 + 1], sh_q[kMaxChi / 2 + 1], sh_any;
    __shared__ double red[kThreads];
    const int b = bonds[blockIdx.x];
    const int dim = bond[b + 1];
    const cd2* L = lenvs + (size_t)(b + 1) * ld * ld;
    cd2* H = work + (size_t)blockIdx.x * ld * ld;
    double* ev = evals_out + (size_t)blockIdx.x * ld;
    for (int idx = threadIdx.x; idx < dim * dim; idx += kThreads) {
        const int r = idx / dim, c = idx - r * dim;
        H[(size_t)r * ld + c] = L[(size_t)r * ld + c];
    }
    __syncthreads();
    wg_hermitian_eigenvalues(H, dim, ld, ev, sh_c, sh_s, sh_e, sh_p, sh_q, &sh_any);
    double lt = 0.0;
    for (int j = 0; j <= b + 1; ++j) lt += log_traces[j];
    // S_k^2 = tr(G) lambda_k; keep S_k > 1e-15, i.e. log lambda_k + lt > log 1e-30
    const double floor_log = -69.07755278982137;      // log(1e-30)
    double part = 0.0;
    for (int k = threadIdx.x; k < dim; k += kThreads) {
        const double l = ev[k];
        if (l > 0.0 && log(l) + lt > floor_log) part += l;
    }
    const double tot = block_sum(part, red);
    double ent = 0.0;
    for (int k = threadIdx.x; k < dim; k += kThreads) {
        const double l = ev[k];
        if (l > 0.0 && log(l) + lt > floor_log) {
            const double p = l / tot;
            ent -= p * log(fmax(p, 1e-300));
        }
    }
    ent = block_sum(ent, red);
    if (threadIdx.x == 0) entropy[blockIdx.x] = ent;
}

int check_chain_args(const void* sites, const int32_t* bond, int64_t stride, int n,
                     int chi_max) {
    QS_REQUIRE(sites != nullptr && bond != nullptr, QS_E_ARG, "null sites or bond array");
    QS_REQUIRE(n >= 1, QS_E_ARG, "chain of %d sites", n);
    QS_REQUIRE(chi_max >= 1 && chi_max <= kMaxChi, QS_E_CHI, "chi_max %d outside [1, %d]",
               chi_max, kMaxChi);
    QS_REQUIRE(stride >= (int64_t)2 * chi_max * chi_max, QS_E_ARG,
               "site_stride %lld < 2*chi_max^2", (long long)stride);
    return QS_OK;
}

Op2 make_op(const double* m) {
    Op2 o;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) o.m[r][c] = mk2(m[2 * (2 * r + c)], m[2 * (2 * r + c) + 1]);
    return o;
}

// One transfer launch over `count` sites starting at site0.
int transfer(const cd2* xs, const int32_t* xb, int64_t xst, const cd2* ys, const int32_t* yb,
             int64_t yst, int site0, int count, const cd2* ein, int64_t ein_step, cd2* eout,
             int64_t eout_step, int ld, const Op2* op, int scale_mode, double* logs,
             hipStream_t hs) {
    const int tiles = (ld + kTile - 1) / kTile;
    Op2 o{};
    hipLaunchKernelGGL(mps_transfer_kernel, dim3(tiles, tiles, count), dim3(kThreads), 0, hs,
                       xs, xb, xst, ys, yb, yst, site0, ein, ein_step, eout, eout_step, ld,
                       op ? *op : o, op ? 1 : 0, scale_mode, logs);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" {

int qs_mps_left_envs(const double* sites, const int32_t* bond, int64_t site_stride, int n,
                     int chi_max, double* envs, size_t env_bytes, double* log_traces,
                     void* stream) {
    clear_error();
    int st = check_chain_args(sites, bond, site_stride, n, chi_max);
    if (st) return st;
    QS_REQUIRE(envs != nullptr, QS_E_ARG, "null environment buffer");
    const size_t need = (size_t)(n + 1) * chi_max * chi_max * sizeof(cd2);
    QS_REQUIRE(env_bytes >= need, QS_E_RANGE, "environment buffer %zu bytes < %zu needed",
               env_bytes, need);
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    const cd2* S = reinterpret_cast<const cd2*>(sites);
    cd2* E = reinterpret_cast<cd2*>(envs);
    const int64_t ld2 = (int64_t)chi_max * chi_max;
    hipLaunchKernelGGL(mps_unit_env_kernel, dim3(1), dim3(64), 0, hs, E);
    QS_HIP(hipGetLastError());
    for (int i = 0; i < n; ++i) {
        st = transfer(S, bond, site_stride, S, bond, site_stride, i, 1, E + i * ld2, 0,
                      E + (i + 1) * ld2, 0, chi_max, nullptr, 1, log_traces, hs);
        if (st) return st;
    }
    hipLaunchKernelGGL(mps_trace_normalise_kernel, dim3(n), dim3(kThreads), 0, hs, bond, 1, n,
                       chi_max, E);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

size_t qs_mps_entropy_work_bytes(int nbonds, int chi_max) {
    if (nbonds < 1 || chi_max < 1) return 0;
    return (size_t)nbonds * chi_max * chi_max * sizeof(cd2) +
           (size_t)nbonds * chi_max * sizeof(double) + (size_t)nbonds * sizeof(int32_t);
}

int qs_mps_entanglement_entropy(const int32_t* bond, int n, int chi_max, const double* lenvs,
                                const double* log_traces, const int32_t* bonds, int nbonds,
                                double* entropy, void* work, size_t work_bytes,
                                void* stream) {
    clear_error();
    QS_REQUIRE(bond && lenvs && log_traces && entropy && work, QS_E_ARG,
               "null pointer argument");
    QS_REQUIRE(n >= 2 && chi_max >= 1 && chi_max <= kMaxChi, QS_E_ARG,
               "chain of %d sites, chi_max %d", n, chi_max);
    QS_REQUIRE(nbonds >= 0 && (nbonds == 0 || bonds), QS_E_ARG, "bad bond list");
    for (int i = 0; i < nbonds; ++i)
        QS_REQUIRE(bonds[i] >= 0 && bonds[i] < n - 1, QS_E_ARG,
                   "bond %d out of range [0, %d]", bonds[i], n - 2);
    if (nbonds == 0) return QS_OK;
    QS_REQUIRE(work_bytes >= qs_mps_entropy_work_bytes(nbonds, chi_max), QS_E_RANGE,
               "work buffer %zu bytes < %zu needed", work_bytes,
               qs_mps_entropy_work_bytes(nbonds, chi_max));
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    cd2* H = static_cast<cd2*>(work);
    double* ev = reinterpret_cast<double*>(H + (size_t)nbonds * chi_max * chi_max);
    int32_t* dbonds = reinterpret_cast<int32_t*>(ev + (size_t)nbonds * chi_max);
    int st = h2d_copy_staged(dbonds, bonds, (size_t)nbonds * sizeof(int32_t), hs);
    if (st) return st;
    hipLaunchKernelGGL(mps_entropy_kernel, dim3(nbonds), dim3(kThreads), 0, hs, bond,
                       reinterpret_cast<const cd2*>(lenvs), chi_max, log_traces, dbonds, H, ev,
                       entropy);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

size_t qs_mps_expect_work_bytes(int chi_max) {
    if (chi_max < 1) return 0;
    return (size_t)2 * kExpectChunk * chi_max * chi_max * sizeof(cd2) +
           (size_t)2 * kExpectChunk * sizeof(cd2);
}

int qs_mps_expect(const double* sites, const int32_t* bond, int64_t site_stride, int n,
                  int chi_max, const double* lenvs, const double* renvs, const double op_a[8],
                  const double op_b[8], int span, double* out, void* work, size_t work_bytes,
                  void* stream) {
    clear_error();
    int st = check_chain_args(sites, bond, site_stride, n, chi_max);
    if (st) return st;
    QS_REQUIRE(lenvs && renvs && out && work && op_a, QS_E_ARG, "null pointer argument");
    QS_REQUIRE(span == 1 || (span == 2 && op_b != nullptr), QS_E_ARG,
               "span %d (1: one-site operator, 2: operator pair with op_b)", span);
    QS_REQUIRE(work_bytes >= qs_mps_expect_work_bytes(chi_max), QS_E_RANGE,
               "work buffer %zu bytes < %zu needed", work_bytes,
               qs_mps_expect_work_bytes(chi_max));
    const int count = n - span + 1;
    if (count <= 0) return QS_OK;
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    const cd2* S = reinterpret_cast<const cd2*>(sites);
    const cd2* L = reinterpret_cast<const cd2*>(lenvs);
    const cd2* R = reinterpret_cast<const cd2*>(renvs);
    const int64_t ld2 = (int64_t)chi_max * chi_max;
    cd2* buf0 = static_cast<cd2*>(work);
    cd2* buf1 = buf0 + (size_t)kExpectChunk * ld2;
    cd2* num = buf1 + (size_t)kExpectChunk * ld2;
    cd2* den = num + kExpectChunk;
    const Op2 oa = make_op(op_a);
    const Op2 ob = span == 2 ? make_op(op_b) : oa;
    for (int c0 = 0; c0 < count; c0 += kExpectChunk) {
        const int k = std::min(kExpectChunk, count - c0);
        for (int pass = 0; pass < 2; ++pass) {
            const Op2* first = pass == 0 ? &oa : nullptr;
            const Op2* second = pass == 0 ? &ob : nullptr;
            st = transfer(S, bond, site_stride, S, bond, site_stride, c0, k, L + c0 * ld2, ld2,
                          buf0, ld2, chi_max, first, 0, nullptr, hs);
            if (st) return st;
            const cd2* m = buf0;
            if (span == 2) {
                st = transfer(S, bond, site_stride, S, bond, site_stride, c0 + 1, k, buf0, ld2,
                              buf1, ld2, chi_max, second, 0, nullptr, hs);
                if (st) return st;
                m = buf1;
            }
            hipLaunchKernelGGL(mps_close_kernel, dim3(k), dim3(kThreads), 0, hs, bond, c0, span,
                               m, ld2, R, chi_max, pass == 0 ? num : den);
            QS_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(mps_ratio_kernel, dim3((k + 255) / 256), dim3(256), 0, hs, num, den,
                           k, out + c0);
        QS_HIP(hipGetLastError());
    }
    return QS_OK;
}

size_t qs_mps_overlap_work_bytes(int n, int chi_max) {
    if (n < 1 || chi_max < 1) return 0;
    return (size_t)2 * chi_max * chi_max * sizeof(cd2) + (size_t)n * sizeof(double);
}

int qs_mps_overlap(const double* bra_sites, const int32_t* bra_bond, int64_t bra_stride,
                   const double* ket_sites, const int32_t* ket_bond, int64_t ket_stride, int n,
                   int chi_max, void* work, size_t work_bytes, double out_host[2],
                   void* stream) {
    clear_error();
    int st = check_chain_args(bra_sites, bra_bond, bra_stride, n, chi_max);
    if (st) return st;
    st = check_chain_args(ket_sites, ket_bond, ket_stride, n, chi_max);
    if (st) return st;
    QS_REQUIRE(work != nullptr && out_host != nullptr, QS_E_ARG, "null work or output");
    QS_REQUIRE(work_bytes >= qs_mps_overlap_work_bytes(n, chi_max), QS_E_RANGE,
               "work buffer %zu bytes < %zu needed", work_bytes,
               qs_mps_overlap_work_bytes(n, chi_max));
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    const int64_t ld2 = (int64_t)chi_max * chi_max;
    cd2* e[2] = {static_cast<cd2*>(work), static_cast<cd2*>(work) + ld2};
    double* logs = reinterpret_cast<double*>(static_cast<cd2*>(work) + 2 * ld2);
    hipLaunchKernelGGL(mps_unit_env_kernel, dim3(1), dim3(64), 0, hs, e[0]);
    QS_HIP(hipGetLastError());
    for (int i = 0; i < n; ++i) {
        st = transfer(reinterpret_cast<const cd2*>(bra_sites), bra_bond, bra_stride,
                      reinterpret_cast<const cd2*>(ket_sites), ket_bond, ket_stride, i, 1,
                      e[i & 1], 0, e[(i + 1) & 1], 0, chi_max, nullptr, 2, logs, hs);
        if (st) return st;
    }
    cd2 last;
    std::vector<double> lg(n);
    QS_HIP(hipMemcpyAsync(&last, e[n & 1], sizeof(cd2), hipMemcpyDeviceToHost, hs));
    QS_HIP(hipMemcpyAsync(lg.data(), logs, n * sizeof(double), hipMemcpyDeviceToHost, hs));
    QS_HIP(hipStreamSynchronize(hs));
    double total = 0.0;
    for (int i = 0; i < n; ++i) total += lg[i];
    const double mag = std::hypot(last.x, last.y);
    out_host[0] = mag > 0.0 ? total + std::log(mag) : -INFINITY;
    out_host[1] = std::atan2(last.y, last.x);
    return QS_OK;
}

}  // extern "C"
