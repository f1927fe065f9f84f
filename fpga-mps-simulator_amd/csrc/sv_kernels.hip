// Statevector kernels for gfx950: fused tile passes, fill, norm, scale, reset.
//
// The reference runs one full-state pass per kernel op (opt_kernel.cpp:71-195:
// dense 2x2 over 2^(n-1) pairs, CX swap and CZ negate over 2^(n-2) indices).
// Here one launch = one HBM pass that applies many gates:
//   * the pass owns a tile of K qubits (qubits 0..3 always among them, so every
//     workgroup streams 256-byte runs); workgroup b owns the 2^K amplitudes that
//     share the bits of b on the other nq-K qubits,
//   * each thread keeps 2^R amplitudes in registers; a "stage" names which R tile
//     qubits are the register qubits, gates on them run in registers, and a change
//     of stage is a transpose through LDS,
//   * diagonal gates and control qubits need no register placement: their bits are
//     read off each amplitude's global index.
#include "qsim_internal.h"

namespace qs {

namespace {

__device__ __forceinline__ cd2 cmul(cd2 a, cd2 b) {
    return mk2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}

__device__ __forceinline__ cd2 cmadd(cd2 acc, cd2 a, cd2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}

__device__ __forceinline__ int lds_slot(int l) { return l + (l >> 4); }

constexpr int kLdsSlots = (1 << SV_TILE_BITS) + ((1 << SV_TILE_BITS) >> 4);

// Per-stage addressing: local index and global index of register j of this thread.
template <int R>
struct Layout {
    int loc_t;                 // local index bits contributed by the thread id
    uint64_t glob_t;           // global bits of base | thread part
    int lreg[R];               // local bit of register bit i
    uint64_t greg[R];          // global bit of register bit i

    __device__ __forceinline__ int loc(int j) const {
        int l = loc_t;
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (j & (1 << i)) l |= lreg[i];
        return l;
    }
    __device__ __forceinline__ uint64_t glob(int j) const {
        uint64_t g = glob_t;
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (j & (1 << i)) g |= greg[i];
        return g;
    }
};

struct Regs {
    int v[4];
};

template <int R>
__device__ __forceinline__ void make_layout(Layout<R>& L, const Regs regs, const DevPass* P,
                                            uint64_t base, int t) {
    int regmask = 0;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        L.lreg[i] = 1 << regs.v[i];
        L.greg[i] = 1ull << P->tq[regs.v[i]];
        regmask |= 1 << regs.v[i];
    }
    // Thread bits fill the non-register local positions in ascending order.
    int loc = 0;
    uint64_t g = base;
    int tb = 0;
    const int K = P->K;
    for (int p = 0; p < K; ++p) {
        if ((regmask >> p) & 1) continue;
        const int bit = (t >> tb) & 1;
        loc |= bit << p;
        g |= (uint64_t)bit << P->tq[p];
        ++tb;
    }
    L.loc_t = loc;
    L.glob_t = g;
}

template <int R, int B>
__device__ __forceinline__ void apply_u1_bit(cd2 (&a)[1 << R], const DevGate& G,
                                             const Layout<R>& L) {
    const cd2 m00 = mk2(G.m[0], G.m[1]), m01 = mk2(G.m[2], G.m[3]);
    const cd2 m10 = mk2(G.m[4], G.m[5]), m11 = mk2(G.m[6], G.m[7]);
    const uint64_t cm = G.cmask;
    if (cm == 0) {
#pragma unroll
        for (int j = 0; j < (1 << R); ++j) {
            if ((j >> B) & 1) continue;           // compile-time after unrolling
            const cd2 x0 = a[j], x1 = a[j | (1 << B)];
            a[j] = cmadd(cmul(m00, x0), m01, x1);
            a[j | (1 << B)] = cmadd(cmul(m10, x0), m11, x1);
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        if ((j >> B) & 1) continue;
        const bool ok = (L.glob(j) & cm) == cm;
        const cd2 x0 = a[j], x1 = a[j | (1 << B)];
        const cd2 y0 = cmadd(cmul(m00, x0), m01, x1);
        const cd2 y1 = cmadd(cmul(m10, x0), m11, x1);
        a[j] = ok ? y0 : x0;
        a[j | (1 << B)] = ok ? y1 : x1;
    }
}

template <int R, int B0, int B1>
__device__ __forceinline__ void apply_u2_bits(cd2 (&a)[1 << R], const DevGate& G,
                                              const Layout<R>& L) {
    const uint64_t cm = G.cmask;
    const cd2* m = reinterpret_cast<const cd2*>(G.m);
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        if (j & ((1 << B0) | (1 << B1))) continue;
        if (cm != 0 && (L.glob(j) & cm) != cm) continue;
        constexpr int b0 = 1 << B0, b1 = 1 << B1;
        const cd2 x0 = a[j], x1 = a[j | b1], x2 = a[j | b0], x3 = a[j | b0 | b1];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const cd2 y = cmadd(cmadd(cmadd(cmul(m[4 * r], x0), m[4 * r + 1], x1),
                                          m[4 * r + 2], x2), m[4 * r + 3], x3);
            if (r == 0) a[j] = y;
            else if (r == 1) a[j | b1] = y;
            else if (r == 2) a[j | b0] = y;
            else a[j | b0 | b1] = y;
        }
    }
}

template <int R>
__device__ __forceinline__ void apply_u1(cd2 (&a)[1 << R], const DevGate& G,
                                         const Layout<R>& L) {
    switch (G.r0) {
        case 0: apply_u1_bit<R, 0>(a, G, L); break;
        case 1: if constexpr (R > 1) apply_u1_bit<R, (R > 1 ? 1 : 0)>(a, G, L); break;
        case 2: if constexpr (R > 2) apply_u1_bit<R, (R > 2 ? 2 : 0)>(a, G, L); break;
        case 3: if constexpr (R > 3) apply_u1_bit<R, (R > 3 ? 3 : 0)>(a, G, L); break;
        default: break;
    }
}

template <int R>
__device__ __forceinline__ void apply_u2(cd2 (&a)[1 << R], const DevGate& G,
                                         const Layout<R>& L) {
    constexpr int S1 = R > 1 ? 1 : 0, S2 = R > 2 ? 2 : 0, S3 = R > 3 ? 3 : 0;
    const int code = G.r0 * 4 + G.r1;
    if constexpr (R >= 2) {
        switch (code) {
            case 0 * 4 + 1: apply_u2_bits<R, 0, S1>(a, G, L); return;
            case 1 * 4 + 0: apply_u2_bits<R, S1, 0>(a, G, L); return;
            default: break;
        }
    }
    if constexpr (R >= 3) {
        switch (code) {
            case 0 * 4 + 2: apply_u2_bits<R, 0, S2>(a, G, L); return;
            case 2 * 4 + 0: apply_u2_bits<R, S2, 0>(a, G, L); return;
            case 1 * 4 + 2: apply_u2_bits<R, S1, S2>(a, G, L); return;
            case 2 * 4 + 1: apply_u2_bits<R, S2, S1>(a, G, L); return;
            default: break;
        }
    }
    if constexpr (R >= 4) {
        switch (code) {
            case 0 * 4 + 3: apply_u2_bits<R, 0, S3>(a, G, L); return;
            case 3 * 4 + 0: apply_u2_bits<R, S3, 0>(a, G, L); return;
            case 1 * 4 + 3: apply_u2_bits<R, S1, S3>(a, G, L); return;
            case 3 * 4 + 1: apply_u2_bits<R, S3, S1>(a, G, L); return;
            case 2 * 4 + 3: apply_u2_bits<R, S2, S3>(a, G, L); return;
            case 3 * 4 + 2: apply_u2_bits<R, S3, S2>(a, G, L); return;
            default: break;
        }
    }
}

template <int R>
__device__ __forceinline__ void apply_diag(cd2 (&a)[1 << R], const DevGate& G,
                                           const Layout<R>& L) {
    const uint64_t cm = G.cmask;
    const int qa = G.qa, qb = G.qb;
    if (G.kind == DG_DIAG1) {
        const cd2 d0 = mk2(G.m[0], G.m[1]), d1 = mk2(G.m[2], G.m[3]);
#pragma unroll
        for (int j = 0; j < (1 << R); ++j) {
            const uint64_t g = L.glob(j);
            if (cm != 0 && (g & cm) != cm) continue;
            a[j] = cmul(((g >> qa) & 1ull) ? d1 : d0, a[j]);
        }
        return;
    }
    // DG_DIAG2: index = 2*bit(qa) + bit(qb)
    const cd2 d00 = mk2(G.m[0], G.m[1]), d01 = mk2(G.m[2], G.m[3]);
    const cd2 d10 = mk2(G.m[4], G.m[5]), d11 = mk2(G.m[6], G.m[7]);
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        const uint64_t g = L.glob(j);
        if (cm != 0 && (g & cm) != cm) continue;
        const bool ba = (g >> qa) & 1ull, bb = (g >> qb) & 1ull;
        const cd2 d = ba ? (bb ? d11 : d10) : (bb ? d01 : d00);
        a[j] = cmul(d, a[j]);
    }
}

// Mask of register indices j (bit j set) whose register bit r is 1.
__device__ __forceinline__ constexpr unsigned reg_col_mask(int r) {
    return r == 0 ? 0xAAAAu : r == 1 ? 0xCCCCu : r == 2 ? 0xF0F0u : 0xFF00u;
}

__device__ __forceinline__ bool is_unit(cd2 z) { return z.x == 1.0 && z.y == 0.0; }

// Applies a unitary diagonal group (DevPhase).  Block-class factors come from the
// prologue's table, thread-class factors are folded per thread, register factors are
// applied to the halves/quarters of the register block they select.
template <int R>
__device__ __forceinline__ void apply_phase(cd2 (&a)[1 << R], const DevPhase& Q,
                                            const DevTerm* __restrict__ terms,
                                            const cd2* __restrict__ bf, int t) {
    constexpr int NR = 1 << R;
    cd2 A = mk2(Q.c[0], Q.c[1]);
    // per-register-bit factors as scalars (an indexed array would live in scratch)
    cd2 k0 = mk2(1.0, 0.0), k1 = k0, k2 = k0, k3 = k0;
    if (bf) {
        A = cmul(A, bf[0]);
        for (int q = 0; q < 8; ++q)
            if ((t >> q) & 1) A = cmul(A, bf[1 + q]);
        k0 = bf[9];
        if (R > 1) k1 = bf[10];
        if (R > 2) k2 = bf[11];
        if (R > 3) k3 = bf[12];
    }
    k0 = cmul(k0, mk2(Q.kr[0][0], Q.kr[0][1]));
    if (R > 1) k1 = cmul(k1, mk2(Q.kr[1][0], Q.kr[1][1]));
    if (R > 2) k2 = cmul(k2, mk2(Q.kr[2][0], Q.kr[2][1]));
    if (R > 3) k3 = cmul(k3, mk2(Q.kr[3][0], Q.kr[3][1]));
    for (int k = 0; k < Q.nt; ++k) {
        const DevTerm& T = terms[Q.t0 + k];
        const cd2 f = mk2(T.re, T.im);
        if (T.cls == PT_T1) {
            if ((t >> T.a) & 1) A = cmul(A, f);
        } else if (T.cls == PT_TT) {
            if (((t >> T.a) & (t >> T.b)) & 1) A = cmul(A, f);
        } else if ((t >> T.b) & 1) {  // PT_TR
            if (T.a == 0) k0 = cmul(k0, f);
            else if (T.a == 1) k1 = cmul(k1, f);
            else if (T.a == 2) k2 = cmul(k2, f);
            else k3 = cmul(k3, f);
        }
    }
    if (Q.sign) {
        // every factor is +-1: the 16-bit parity mask of the register block
        unsigned M = (A.x < 0.0) ? 0xFFFFu : 0u;
        if (k0.x < 0.0) M ^= reg_col_mask(0);
        if (R > 1 && k1.x < 0.0) M ^= reg_col_mask(1);
        if (R > 2 && k2.x < 0.0) M ^= reg_col_mask(2);
        if (R > 3 && k3.x < 0.0) M ^= reg_col_mask(3);
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int s2 = r + 1; s2 < R; ++s2) {
                const int pi = r == 0 ? s2 - 1 : (r == 1 ? s2 + 1 : 5);
                if (Q.crr[pi][0] < 0.0) M ^= reg_col_mask(r) & reg_col_mask(s2);
            }
#pragma unroll
        for (int j = 0; j < NR; ++j)
            if ((M >> j) & 1u) a[j] = -a[j];
        return;
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        cd2 f = A;
        if (j & 1) f = cmul(f, k0);
        if (R > 1 && (j & 2)) f = cmul(f, k1);
        if (R > 2 && (j & 4)) f = cmul(f, k2);
        if (R > 3 && (j & 8)) f = cmul(f, k3);
        a[j] = cmul(f, a[j]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int s2 = r + 1; s2 < R; ++s2) {
            const int pi = r == 0 ? s2 - 1 : (r == 1 ? s2 + 1 : 5);
            const cd2 f = mk2(Q.crr[pi][0], Q.crr[pi][1]);
            if (!is_unit(f)) {
#pragma unroll
                for (int j = 0; j < NR; ++j)
                    if ((j >> r) & (j >> s2) & 1) a[j] = cmul(f, a[j]);
            }
        }
}

// Block-class phase factors of every phase group of one pass, one thread per
// (slot, workgroup): [A_B, lambda_0..7, kappa_0..3].
__global__ void sv_phase_block_kernel(int nq, const DevPass* __restrict__ P,
                                      const DevPhase* __restrict__ phases,
                                      const DevTerm* __restrict__ terms,
                                      cd2* __restrict__ factors, uint32_t nblocks) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nslot = P->nslot;
    if (gid >= (uint64_t)nslot * nblocks) return;
    const int slot = (int)(gid / nblocks);
    const uint32_t blk = (uint32_t)(gid - (uint64_t)slot * nblocks);
    const DevPhase* Q = nullptr;
    for (int k = 0; k < P->nphase; ++k)
        if (phases[P->phase0 + k].slot == slot) Q = &phases[P->phase0 + k];
    uint64_t base = 0;
    {
        const uint64_t tmask = P->tile_mask;
        uint64_t b = blk;
        for (int q = 0; q < nq && b; ++q) {
            if ((tmask >> q) & 1ull) continue;
            base |= (b & 1ull) << q;
            b >>= 1;
        }
    }
    cd2 v[PHASE_BLOCK_VALUES];
#pragma unroll
    for (int k = 0; k < PHASE_BLOCK_VALUES; ++k) v[k] = mk2(1.0, 0.0);
    const DevTerm* tm = terms + P->term0 + Q->b0;
    for (int k = 0; k < Q->nb; ++k) {
        const DevTerm& T = tm[k];
        const cd2 f = mk2(T.re, T.im);
        if (T.cls == PT_BB) {
            const bool on = ((base >> T.a) & 1ull) && (T.b < 0 || ((base >> T.b) & 1ull));
            if (on) v[0] = cmul(v[0], f);
        } else if ((base >> T.b) & 1ull) {
            const int idx = (T.cls == PT_BT) ? 1 + T.a : 9 + T.a;
#pragma unroll
            for (int i = 1; i < PHASE_BLOCK_VALUES; ++i)
                if (i == idx) v[i] = cmul(v[i], f);
        }
    }
    cd2* out = factors + ((size_t)(P->slot0 + slot) * nblocks + blk) * PHASE_BLOCK_VALUES;
#pragma unroll
    for (int k = 0; k < PHASE_BLOCK_VALUES; ++k) out[k] = v[k];
}

template <int R>
__global__ void __launch_bounds__(1 << (SV_TILE_BITS - SV_REG_BITS))
sv_pass_kernel(cd2* __restrict__ state, int nq, const DevPass* __restrict__ P,
               const DevStage* __restrict__ stages, const DevGate* __restrict__ gates,
               const DevPhase* __restrict__ phases, const DevTerm* __restrict__ terms,
               const cd2* __restrict__ factors) {
    constexpr int NR = 1 << R;
    __shared__ cd2 lds[kLdsSlots];

    const int t = threadIdx.x;
    const int K = P->K;
    const int nstage = P->nstage;
    const DevStage* st = stages + P->stage0;
    const DevGate* gt = gates + P->gate0;
    const DevPhase* ph = phases + P->phase0;
    const DevTerm* tm = terms + P->term0;
    const uint32_t nblocks = gridDim.x;
    // Block base: bits of blockIdx.x deposited into the non-tile qubits, ascending.
    uint64_t base = 0;
    if (!P->tile_is_all) {
        const uint64_t tmask = P->tile_mask;
        uint64_t b = blockIdx.x;
        for (int q = 0; q < nq && b; ++q) {
            if ((tmask >> q) & 1ull) continue;
            base |= (b & 1ull) << q;
            b >>= 1;
        }
    }

    // Load layout: register bits are the top R local positions; thread bits are the
    // low ones, so lanes walk consecutive addresses on qubits 0..3.
    Regs load_regs{{0, 0, 0, 0}};
#pragma unroll
    for (int i = 0; i < R; ++i) load_regs.v[i] = K - R + i;
    Layout<R> Lcur;
    make_layout<R>(Lcur, load_regs, P, base, t);
    Regs cur = load_regs;

    cd2 a[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) a[j] = state[Lcur.glob(j)];

    for (int s = 0; s < nstage; ++s) {
        Regs want{{0, 0, 0, 0}};
        bool same = true;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            want.v[i] = st[s].regs[i];
            same = same && (want.v[i] == cur.v[i]);
        }
        if (!same) {
#pragma unroll
            for (int j = 0; j < NR; ++j) lds[lds_slot(Lcur.loc(j))] = a[j];
            __syncthreads();
            make_layout<R>(Lcur, want, P, base, t);
            cur = want;
#pragma unroll
            for (int j = 0; j < NR; ++j) a[j] = lds[lds_slot(Lcur.loc(j))];
            __syncthreads();
        }
        const int g1 = st[s].g1;
        for (int gi = st[s].g0; gi < g1; ++gi) {
            const DevGate& G = gt[gi];
            switch (G.kind) {
                case DG_U1: apply_u1<R>(a, G, Lcur); break;
                case DG_U2: apply_u2<R>(a, G, Lcur); break;
                case DG_PHASE: {
                    const DevPhase& Q = ph[G.qa];
                    const cd2* bf =
                        Q.slot >= 0 ? factors + ((size_t)(P->slot0 + Q.slot) * nblocks + blockIdx.x) *
                                                    PHASE_BLOCK_VALUES
                                    : nullptr;
                    apply_phase<R>(a, Q, tm, bf, t);
                    break;
                }
                default: apply_diag<R>(a, G, Lcur); break;
            }
        }
    }

    bool back = false;
#pragma unroll
    for (int i = 0; i < R; ++i) back = back || (cur.v[i] != load_regs.v[i]);
    if (back) {
#pragma unroll
        for (int j = 0; j < NR; ++j) lds[lds_slot(Lcur.loc(j))] = a[j];
        __syncthreads();
        make_layout<R>(Lcur, load_regs, P, base, t);
#pragma unroll
        for (int j = 0; j < NR; ++j) a[j] = lds[lds_slot(Lcur.loc(j))];
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) state[Lcur.glob(j)] = a[j];
}

__global__ void sv_fill_zero_kernel(cd2* __restrict__ state, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        state[i] = mk2(i == 0 ? 1.0 : 0.0, 0.0);
}

__global__ void sv_scale_kernel(cd2* __restrict__ state, uint64_t n, cd2 f) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        state[i] = cmul(f, state[i]);
}

// Zero the amplitudes whose bit q is set.
__global__ void sv_project_kernel(cd2* __restrict__ state, uint64_t n, int q) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        if ((i >> q) & 1ull) state[i] = mk2(0.0, 0.0);
}

__global__ void sv_probs_kernel(const cd2* __restrict__ state, uint64_t n,
                                double* __restrict__ p) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const cd2 v = state[i];
        p[i] = fma(v.x, v.x, v.y * v.y);
    }
}

constexpr int kNormBlocks = 1024;
constexpr int kNormThreads = 256;

// Per-block partial sums of |a|^2 (fixed grid, fixed order: deterministic).
__global__ void __launch_bounds__(kNormThreads)
sv_norm_partial_kernel(const cd2* __restrict__ state, uint64_t n, double* __restrict__ part) {
    __shared__ double red[kNormThreads];
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const cd2 v = state[i];
        acc = fma(v.x, v.x, fma(v.y, v.y, acc));
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kNormThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

int grid_for(uint64_t n, int threads) {
    uint64_t b = (n + threads - 1) / threads;
    if (b > 65536) b = 65536;
    return (int)(b == 0 ? 1 : b);
}

}  // namespace

int sv_launch_plan(double2* state, int nq, const SvPlan& plan, const void* dev_blob,
                   hipStream_t stream) {
    const char* blob = reinterpret_cast<const char*>(dev_blob);
    const SvBlobLayout L = sv_blob_layout(plan, nq);
    const DevPass* dpasses = reinterpret_cast<const DevPass*>(blob + L.passes);
    const DevStage* dstages = reinterpret_cast<const DevStage*>(blob + L.stages);
    const DevGate* dgates = reinterpret_cast<const DevGate*>(blob + L.gates);
    const DevPhase* dphases = reinterpret_cast<const DevPhase*>(blob + L.phases);
    const DevTerm* dterms = reinterpret_cast<const DevTerm*>(blob + L.terms);
    double2* dfactors = reinterpret_cast<double2*>(const_cast<char*>(blob) + L.factors);
    for (size_t p = 0; p < plan.passes.size(); ++p) {
        const DevPass& P = plan.passes[p];
        const DevPass* dP = dpasses + p;
        const unsigned blocks = P.tile_is_all ? 1u : (1u << (nq - P.K));
        const unsigned threads = 1u << (P.K - P.R);
        if (P.nslot > 0) {
            const uint64_t work = (uint64_t)P.nslot * blocks;
            hipLaunchKernelGGL(sv_phase_block_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256),
                               0, stream, nq, dP, dphases, dterms, reinterpret_cast<cd2*>(dfactors), blocks);
            QS_HIP(hipGetLastError());
        }
        switch (P.R) {
            case 1: hipLaunchKernelGGL(sv_pass_kernel<1>, dim3(blocks), dim3(threads), 0, stream,
                                       reinterpret_cast<cd2*>(state), nq, dP, dstages, dgates, dphases, dterms, reinterpret_cast<cd2*>(dfactors)); break;
            case 2: hipLaunchKernelGGL(sv_pass_kernel<2>, dim3(blocks), dim3(threads), 0, stream,
                                       reinterpret_cast<cd2*>(state), nq, dP, dstages, dgates, dphases, dterms, reinterpret_cast<cd2*>(dfactors)); break;
            case 3: hipLaunchKernelGGL(sv_pass_kernel<3>, dim3(blocks), dim3(threads), 0, stream,
                                       reinterpret_cast<cd2*>(state), nq, dP, dstages, dgates, dphases, dterms, reinterpret_cast<cd2*>(dfactors)); break;
            default: hipLaunchKernelGGL(sv_pass_kernel<4>, dim3(blocks), dim3(threads), 0, stream,
                                        reinterpret_cast<cd2*>(state), nq, dP, dstages, dgates, dphases, dterms, reinterpret_cast<cd2*>(dfactors)); break;
        }
        QS_HIP(hipGetLastError());
    }
    return QS_OK;
}

int sv_fill_zero(double2* state, int nq, hipStream_t stream) {
    const uint64_t n = 1ull << nq;
    hipLaunchKernelGGL(sv_fill_zero_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, reinterpret_cast<cd2*>(state), n);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

int sv_scale(double2* state, int nq, double re, double im, hipStream_t stream) {
    const uint64_t n = 1ull << nq;
    hipLaunchKernelGGL(sv_scale_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream,
                       reinterpret_cast<cd2*>(state), n, mk2(re, im));
    QS_HIP(hipGetLastError());
    return QS_OK;
}

int sv_project(double2* state, int nq, int q, hipStream_t stream) {
    const uint64_t n = 1ull << nq;
    hipLaunchKernelGGL(sv_project_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, reinterpret_cast<cd2*>(state), n, q);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

int sv_probs(const double2* state, int nq, double* p, hipStream_t stream) {
    const uint64_t n = 1ull << nq;
    hipLaunchKernelGGL(sv_probs_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, reinterpret_cast<const cd2*>(state), n, p);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

int sv_norm_sq(const double2* state, int nq, double* out, hipStream_t stream) {
    const uint64_t n = 1ull << nq;
    double* part = nullptr;
    QS_HIP(hipMallocAsync(reinterpret_cast<void**>(&part), kNormBlocks * sizeof(double), stream));
    hipLaunchKernelGGL(sv_norm_partial_kernel, dim3(kNormBlocks), dim3(kNormThreads), 0, stream,
                       reinterpret_cast<const cd2*>(state), n, part);
    QS_HIP(hipGetLastError());
    std::vector<double> h(kNormBlocks);
    QS_HIP(hipMemcpyAsync(h.data(), part, kNormBlocks * sizeof(double), hipMemcpyDeviceToHost,
                          stream));
    QS_HIP(hipFreeAsync(part, stream));
    QS_HIP(hipStreamSynchronize(stream));
    double s = 0.0;
    for (double v : h) s += v;
    *out = s;
    return QS_OK;
}

}  // namespace qs
