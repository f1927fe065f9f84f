// Device building blocks of the statevector pass kernels: register layouts, the op
// bodies on a thread's register block, the phase groups and the LDS transpose, and
// op_body, which runs one op of a pass.  Included by sv_kernels.hip (the op-stream
// interpreter) and handed as it is to the run-time compiler for the per-plan kernels
// of sv_jit.cpp, so both run the same code.
#pragma once

#include "sv_device_types.h"

namespace qs {

__device__ __forceinline__ cd2 cmul(cd2 a, cd2 b) {
    return mk2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}

__device__ __forceinline__ cd2 cmadd(cd2 acc, cd2 a, cd2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}

__device__ __forceinline__ int lds_slot(int l) { return l + (l >> 4); }

// Pointer into memory that no thread writes during the launch (address space 4): uniform
// loads through it become scalar loads.  Through a plain pointer the compiler may keep
// them on the vector path, where they queue behind the tile's HBM loads and stores.
typedef const cd2 __attribute__((address_space(4)))* const_cd2_ptr;

constexpr int kLdsSlots = (1 << SV_TILE_BITS) + ((1 << SV_TILE_BITS) >> 4);

// Addressing under one register layout: local (LDS) index and the thread's part of
// the global index of register j; the tile's block bits are OR-ed in separately.
template <int R>
struct Layout {
    int loc_t;                 // local index bits contributed by the thread id
    uint64_t glob_t;           // global index bits contributed by the thread id
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

template <int R>
__device__ __forceinline__ void make_layout(Layout<R>& L, const DevLayout& D, int t) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
        L.lreg[i] = 1 << D.regs[i];
        L.greg[i] = 1ull << D.rq[i];
    }
    int loc = 0;
    uint64_t g = 0;
    // thread bits beyond the tile's thread count are 0 (a narrow tile, R < 4, has one
    // thread)
#pragma unroll
    for (int b = 0; b < SV_THREAD_BITS; ++b) {
        const int bit = (t >> b) & 1;
        loc |= bit << D.tpos[b];
        g |= (uint64_t)bit << D.tq[b];
    }
    L.loc_t = loc;
    L.glob_t = g;
}

// Leading part of a DevGate record (kind through m[11]): everything DG_ROT, DG_ROTS,
// DG_PHASE_SIGN (its SignDesc), DG_PHASE (table entry, slot, register-pair factors) and
// DG_SCALE read.
struct alignas(16) OpHead {
    int32_t kind, r0, r1, qa, qb, aux;
    uint64_t cmask;
    double m[12];
};
static_assert(sizeof(OpHead) == 128 && __builtin_offsetof(DevGate, m) == __builtin_offsetof(OpHead, m),
              "OpHead must be the leading part of DevGate");
static_assert(sizeof(SignDesc) <= sizeof(double) * 12, "SignDesc must lie in the op head");
static_assert(__builtin_offsetof(SignDesc, mb) == 4 && __builtin_offsetof(SignDesc, pp) == 4 * (1 + SV_THREAD_BITS) &&
                  __builtin_offsetof(SignDesc, slot) == 4 * (1 + 2 * SV_THREAD_BITS),
              "SignDesc is read as consecutive 32-bit words");

// Global index bits of tile `tile`: bit i of the tile index lands on block qubit bq[i].
// Unrolled over the pass's runs of consecutive block qubits so that the run table
// arrives in one batch of scalar loads.
__device__ __forceinline__ uint64_t tile_base(const DevPass* P, uint32_t tile) {
    uint64_t base = 0;
#pragma unroll
    for (int k = 0; k < SV_TILE_BITS + 1; ++k)
        base |= (uint64_t)((tile >> P->run_src[k]) & P->run_mask[k]) << P->run_dst[k];
    return base;
}

// Uncontrolled 2x2 matrix m (row-major complex) on register bit B, in place.
template <int R, int B>
__device__ __forceinline__ void u1_bit(cd2 (&a)[1 << R], const double* __restrict__ m) {
    const cd2 m00 = mk2(m[0], m[1]), m01 = mk2(m[2], m[3]);
    const cd2 m10 = mk2(m[4], m[5]), m11 = mk2(m[6], m[7]);
    int npair = 0;
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        if ((j >> B) & 1) continue;           // compile-time after unrolling
        const cd2 x0 = a[j], x1 = a[j | (1 << B)];
        a[j] = cmadd(cmul(m00, x0), m01, x1);
        a[j | (1 << B)] = cmadd(cmul(m10, x0), m11, x1);
        // Two pairs (eight independent FMA chains) per scheduling group keep the FP64
        // pipe fed without holding every pair's results in temporaries at once.
        if (++npair % 2 == 0) __builtin_amdgcn_sched_barrier(0);
    }
}

// Controlled 2x2 matrix on register bit B (DG_U1): amplitudes whose global index
// lacks a control bit keep their value (branch-free select).
template <int R, int B>
__device__ __forceinline__ void apply_cu1_bit(cd2 (&a)[1 << R], uint64_t cm,
                                              const double* __restrict__ m,
                                              const Layout<R>& L, uint64_t base) {
    const cd2 m00 = mk2(m[0], m[1]), m01 = mk2(m[2], m[3]);
    const cd2 m10 = mk2(m[4], m[5]), m11 = mk2(m[6], m[7]);
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        if ((j >> B) & 1) continue;
        const bool ok = ((base | L.glob(j)) & cm) == cm;
        const cd2 x0 = a[j], x1 = a[j | (1 << B)];
        const cd2 y0 = cmadd(cmul(m00, x0), m01, x1);
        const cd2 y1 = cmadd(cmul(m10, x0), m11, x1);
        a[j] = ok ? y0 : x0;
        a[j | (1 << B)] = ok ? y1 : x1;
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Two-qubit matrix on register bits B0 (row-index MSB) and B1; amplitudes whose global
// index lacks a control bit keep their value (branch-free select; cmask 0 selects all).
template <int R, int B0, int B1>
__device__ __forceinline__ void apply_u2_bits(cd2 (&a)[1 << R], uint64_t cm,
                                              const double* __restrict__ md,
                                              const Layout<R>& L, uint64_t base) {
    const cd2* m = reinterpret_cast<const cd2*>(md);
    constexpr int b0 = 1 << B0, b1 = 1 << B1;
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        if (j & (b0 | b1)) continue;          // compile-time after unrolling
        const bool ok = ((base | L.glob(j)) & cm) == cm;
        const cd2 x0 = a[j], x1 = a[j | b1], x2 = a[j | b0], x3 = a[j | b0 | b1];
        cd2 y[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            y[r] = cmadd(cmadd(cmadd(cmul(m[4 * r], x0), m[4 * r + 1], x1), m[4 * r + 2], x2),
                         m[4 * r + 3], x3);
        a[j] = ok ? y[0] : x0;
        a[j | b1] = ok ? y[1] : x1;
        a[j | b0] = ok ? y[2] : x2;
        a[j | b0 | b1] = ok ? y[3] : x3;
        __builtin_amdgcn_sched_barrier(0);
    }
}

// a * b + c into a register of its own (three-address v_fma_f64).  Written as plain
// fma, the compiler picks the tied two-address v_fmac_f64 and then copies c when c is
// still needed, which costs a 64-bit move per result.  a is uniform (an SGPR).
__device__ __forceinline__ double fma3(double a, double b, double c) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "s"(a), "v"(b), "v"(c));
    return r;
}

// Scaled rotation on register bit B (DG_ROT).  Form A (FB false) maps a pair to
// (x0 - t x1, x1 + t x0), form B to (t x0 - x1, x0 + t x1); with DIAG the second
// output is multiplied by k.  Four FMAs per pair, plus one complex product with DIAG.
template <int R, int B, bool FB, bool DIAG>
__device__ __forceinline__ void rot_bit(cd2 (&a)[1 << R], double t, cd2 k) {
    int npair = 0;
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        if ((j >> B) & 1) continue;           // compile-time after unrolling
        const cd2 x0 = a[j], x1 = a[j | (1 << B)];
        cd2 y0, y1;
        if constexpr (!FB) {
            y0 = mk2(fma(-t, x1.x, x0.x), fma(-t, x1.y, x0.y));
            y1 = mk2(fma3(t, x0.x, x1.x), fma3(t, x0.y, x1.y));
        } else {
            y0 = mk2(fma(t, x0.x, -x1.x), fma(t, x0.y, -x1.y));
            y1 = mk2(fma3(t, x1.x, x0.x), fma3(t, x1.y, x0.y));
        }
        if constexpr (DIAG) y1 = cmul(k, y1);
        a[j] = y0;
        a[j | (1 << B)] = y1;
        if (++npair % 4 == 0) __builtin_amdgcn_sched_barrier(0);
    }
}

// Ops that differ only in which registers they touch are dispatched by a switch whose
// case bodies are bracketed by distinct empty asm statements: without them the
// optimizer merges the isomorphic bodies behind pointer selects, which moves the
// register block to scratch memory.  Every case body is straight-line code.
#define QS_CASE(tag) asm volatile("; " tag)

template <int R>
__device__ __forceinline__ void apply_u1s(cd2 (&a)[1 << R], int bit, const double* m) {
    constexpr int S1 = R > 1 ? 1 : 0, S2 = R > 2 ? 2 : 0, S3 = R > 3 ? 3 : 0;
    switch (bit) {
        case 0: QS_CASE("u1s 0 <"); u1_bit<R, 0>(a, m); QS_CASE("u1s 0 >"); break;
        case 1: if constexpr (R > 1) { QS_CASE("u1s 1 <"); u1_bit<R, S1>(a, m); QS_CASE("u1s 1 >"); } break;
        case 2: if constexpr (R > 2) { QS_CASE("u1s 2 <"); u1_bit<R, S2>(a, m); QS_CASE("u1s 2 >"); } break;
        default: if constexpr (R > 3) { QS_CASE("u1s 3 <"); u1_bit<R, S3>(a, m); QS_CASE("u1s 3 >"); } break;
    }
}

// DG_ROT + 4 bit + variant: one switch over the sixteen bodies.
#define QS_ROT_CASES(B, TAG)                                                              \
    case DG_ROT + 4 * B + 0: QS_CASE("rot " TAG "a <"); rot_bit<R, B, false, false>(a, t, k); QS_CASE("rot " TAG "a >"); break; \
    case DG_ROT + 4 * B + 1: QS_CASE("rot " TAG "b <"); rot_bit<R, B, true, false>(a, t, k); QS_CASE("rot " TAG "b >"); break;  \
    case DG_ROT + 4 * B + 2: QS_CASE("rot " TAG "c <"); rot_bit<R, B, false, true>(a, t, k); QS_CASE("rot " TAG "c >"); break;  \
    case DG_ROT + 4 * B + 3: QS_CASE("rot " TAG "d <"); rot_bit<R, B, true, true>(a, t, k); QS_CASE("rot " TAG "d >"); break;

template <int R>
__device__ __forceinline__ void apply_rot(cd2 (&a)[1 << R], int kind, double t, cd2 k) {
    if constexpr (R == 4) {
        switch (kind) {
            QS_ROT_CASES(0, "0") QS_ROT_CASES(1, "1") QS_ROT_CASES(2, "2")
            case DG_ROT + 12: QS_CASE("rot 3a <"); rot_bit<R, 3, false, false>(a, t, k); QS_CASE("rot 3a >"); break;
            case DG_ROT + 13: QS_CASE("rot 3b <"); rot_bit<R, 3, true, false>(a, t, k); QS_CASE("rot 3b >"); break;
            case DG_ROT + 14: QS_CASE("rot 3c <"); rot_bit<R, 3, false, true>(a, t, k); QS_CASE("rot 3c >"); break;
            default: QS_CASE("rot 3d <"); rot_bit<R, 3, true, true>(a, t, k); QS_CASE("rot 3d >"); break;
        }
    } else {
        // narrow tiles (fewer than four register bits): general 2x2 on the bit
        const int r1 = (kind - DG_ROT) & 3;
        double m[8] = {1.0, 0.0, -t, 0.0, t, 0.0, 1.0, 0.0};
        if (r1 & 1) {
            m[0] = t; m[2] = -1.0; m[4] = 1.0; m[6] = t;
        }
        if (r1 & 2) {
            const cd2 r0 = cmul(k, mk2(m[4], m[5])), r1v = cmul(k, mk2(m[6], m[7]));
            m[4] = r0.x; m[5] = r0.y; m[6] = r1v.x; m[7] = r1v.y;
        }
        apply_u1s<R>(a, (kind - DG_ROT) >> 2, m);
    }
}
#undef QS_ROT_CASES

// First-form rotations on the register bits of M without their output factors, then
// the factors as one diagonal: amplitude j times f_j (f[2j - 2], f[2j - 1]) when j has a
// bit of M.  The rotations leave the j-th amplitude's phase factor to the diagonal,
// which commutes with them (distinct bits).
template <int R, int M>
__device__ __forceinline__ void rots_body(cd2 (&a)[1 << R], const double* t,
                                          const double* __restrict__ f) {
    const cd2 one = mk2(1.0, 0.0);
    if constexpr ((M & 1) != 0) rot_bit<R, 0, false, false>(a, t[0], one);
    if constexpr ((M & 2) != 0 && R > 1) rot_bit<R, (R > 1 ? 1 : 0), false, false>(a, t[1], one);
    if constexpr ((M & 4) != 0 && R > 2) rot_bit<R, (R > 2 ? 2 : 0), false, false>(a, t[2], one);
    if constexpr ((M & 8) != 0 && R > 3) rot_bit<R, (R > 3 ? 3 : 0), false, false>(a, t[3], one);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 1; j < (1 << R); ++j) {
        if ((j & M) == 0) continue;            // compile-time after unrolling
        a[j] = cmul(mk2(f[2 * j - 2], f[2 * j - 1]), a[j]);
        if (j % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    }
}

// A DG_ROTS op followed by its register-only sign mask S (the op's aux word), both known
// when a per-plan kernel is compiled.  Amplitude j with a merged output factor takes the
// sign on the factor, cmul(-f_j, a_j) = -cmul(f_j, a_j) exactly (up to the sign of a zero
// result), so the negation is a free modifier of the factor's scalar operands.  Left to
// itself the compiler keeps -a_j pending and, where the next op's three-address FMAs or a
// transpose cannot take a modifier, materialises it with two instructions per double.
// Amplitudes without a factor are multiplied by -1.0 from an SGPR the compiler cannot
// fold, one instruction per double, as the interpreter's apply_sign_reg does.
template <int R, int M, unsigned S>
__device__ __forceinline__ void rots_body_signed(cd2 (&a)[1 << R], const double* t,
                                                 const double* __restrict__ f) {
    const cd2 one = mk2(1.0, 0.0);
    if constexpr ((M & 1) != 0) rot_bit<R, 0, false, false>(a, t[0], one);
    if constexpr ((M & 2) != 0 && R > 1) rot_bit<R, (R > 1 ? 1 : 0), false, false>(a, t[1], one);
    if constexpr ((M & 4) != 0 && R > 2) rot_bit<R, (R > 2 ? 2 : 0), false, false>(a, t[2], one);
    if constexpr ((M & 8) != 0 && R > 3) rot_bit<R, (R > 3 ? 3 : 0), false, false>(a, t[3], one);
    __builtin_amdgcn_sched_barrier(0);
    double m1;
    asm("s_mov_b64 %0, -1.0" : "=s"(m1));
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        const bool neg = (S >> j) & 1u;
        if ((j & M) != 0) {
            const cd2 fj = mk2(f[2 * j - 2], f[2 * j - 1]);
            a[j] = cmul(neg ? mk2(-fj.x, -fj.y) : fj, a[j]);
        } else if (neg) {
            a[j] = a[j] * m1;
        }
        if (j % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    }
}

#define QS_ROTS_CASE(M) \
    case DG_ROTS + M - 1: QS_CASE("rots " #M " <"); rots_body<R, M>(a, t, f); QS_CASE("rots " #M " >"); break;

// DG_ROTS + mask - 1: t = m[0..3] from the op head, the merged output factors f_j from
// the record (m[2 + 2j], m[3 + 2j]).
template <int R>
__device__ __forceinline__ void apply_rots(cd2 (&a)[1 << R], int kind, const double* head,
                                           const double* __restrict__ m) {
    const double* t = head;
    const double* f = m + 4;
    if constexpr (R == 4) {
        switch (kind) {
            QS_ROTS_CASE(1) QS_ROTS_CASE(2) QS_ROTS_CASE(3) QS_ROTS_CASE(4) QS_ROTS_CASE(5)
            QS_ROTS_CASE(6) QS_ROTS_CASE(7) QS_ROTS_CASE(8) QS_ROTS_CASE(9) QS_ROTS_CASE(10)
            QS_ROTS_CASE(11) QS_ROTS_CASE(12) QS_ROTS_CASE(13) QS_ROTS_CASE(14)
            default: QS_CASE("rots 15 <"); rots_body<R, 15>(a, t, f); QS_CASE("rots 15 >"); break;
        }
    } else {
        // narrow tiles: one rotation with its own output factor k_b = f_(2^b) per bit
        const int mask = kind - DG_ROTS + 1;
        for (int b = 0; b < R; ++b)
            if ((mask >> b) & 1) {
                const int j = 1 << b;
                apply_rot<R>(a, DG_ROT + 4 * b + 2, t[b], mk2(f[2 * j - 2], f[2 * j - 1]));
            }
    }
}
#undef QS_ROTS_CASE

template <int R>
__device__ __forceinline__ void apply_cu1(cd2 (&a)[1 << R], int bit, uint64_t cm,
                                          const double* m, const Layout<R>& L, uint64_t base) {
    constexpr int S1 = R > 1 ? 1 : 0, S2 = R > 2 ? 2 : 0, S3 = R > 3 ? 3 : 0;
    switch (bit) {
        case 0: QS_CASE("cu1 0 <"); apply_cu1_bit<R, 0>(a, cm, m, L, base); QS_CASE("cu1 0 >"); break;
        case 1: if constexpr (R > 1) { QS_CASE("cu1 1 <"); apply_cu1_bit<R, S1>(a, cm, m, L, base); QS_CASE("cu1 1 >"); } break;
        case 2: if constexpr (R > 2) { QS_CASE("cu1 2 <"); apply_cu1_bit<R, S2>(a, cm, m, L, base); QS_CASE("cu1 2 >"); } break;
        default: if constexpr (R > 3) { QS_CASE("cu1 3 <"); apply_cu1_bit<R, S3>(a, cm, m, L, base); QS_CASE("cu1 3 >"); } break;
    }
}

template <int R>
__device__ __forceinline__ void apply_u2(cd2 (&a)[1 << R], int b0, int b1, uint64_t cm,
                                         const double* m, const Layout<R>& L, uint64_t base) {
    constexpr int S1 = R > 1 ? 1 : 0, S2 = R > 2 ? 2 : 0, S3 = R > 3 ? 3 : 0;
    switch (b0 * 4 + b1) {
        case 0 * 4 + 1: if constexpr (R >= 2) { QS_CASE("u2 01 <"); apply_u2_bits<R, 0, S1>(a, cm, m, L, base); QS_CASE("u2 01 >"); } break;
        case 1 * 4 + 0: if constexpr (R >= 2) { QS_CASE("u2 10 <"); apply_u2_bits<R, S1, 0>(a, cm, m, L, base); QS_CASE("u2 10 >"); } break;
        case 0 * 4 + 2: if constexpr (R >= 3) { QS_CASE("u2 02 <"); apply_u2_bits<R, 0, S2>(a, cm, m, L, base); QS_CASE("u2 02 >"); } break;
        case 2 * 4 + 0: if constexpr (R >= 3) { QS_CASE("u2 20 <"); apply_u2_bits<R, S2, 0>(a, cm, m, L, base); QS_CASE("u2 20 >"); } break;
        case 1 * 4 + 2: if constexpr (R >= 3) { QS_CASE("u2 12 <"); apply_u2_bits<R, S1, S2>(a, cm, m, L, base); QS_CASE("u2 12 >"); } break;
        case 2 * 4 + 1: if constexpr (R >= 3) { QS_CASE("u2 21 <"); apply_u2_bits<R, S2, S1>(a, cm, m, L, base); QS_CASE("u2 21 >"); } break;
        case 0 * 4 + 3: if constexpr (R >= 4) { QS_CASE("u2 03 <"); apply_u2_bits<R, 0, S3>(a, cm, m, L, base); QS_CASE("u2 03 >"); } break;
        case 3 * 4 + 0: if constexpr (R >= 4) { QS_CASE("u2 30 <"); apply_u2_bits<R, S3, 0>(a, cm, m, L, base); QS_CASE("u2 30 >"); } break;
        case 1 * 4 + 3: if constexpr (R >= 4) { QS_CASE("u2 13 <"); apply_u2_bits<R, S1, S3>(a, cm, m, L, base); QS_CASE("u2 13 >"); } break;
        case 3 * 4 + 1: if constexpr (R >= 4) { QS_CASE("u2 31 <"); apply_u2_bits<R, S3, S1>(a, cm, m, L, base); QS_CASE("u2 31 >"); } break;
        case 2 * 4 + 3: if constexpr (R >= 4) { QS_CASE("u2 23 <"); apply_u2_bits<R, S2, S3>(a, cm, m, L, base); QS_CASE("u2 23 >"); } break;
        default: if constexpr (R >= 4) { QS_CASE("u2 32 <"); apply_u2_bits<R, S3, S2>(a, cm, m, L, base); QS_CASE("u2 32 >"); } break;
    }
}

// Uncontrolled SWAP of register bits B0 and B1: a permutation of the register block.
template <int R, int B0, int B1>
__device__ __forceinline__ void swap_bits(cd2 (&a)[1 << R]) {
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        if (((j >> B0) & 1) && !((j >> B1) & 1)) {   // compile-time after unrolling
            const int k = j ^ (1 << B0) ^ (1 << B1);
            const cd2 x = a[j];
            a[j] = a[k];
            a[k] = x;
        }
    }
}

// DG_SWAP: r0 indexes the register-bit pairs (0,1) (0,2) (0,3) (1,2) (1,3) (2,3);
// the planner emits it for four-bit register blocks only.
template <int R>
__device__ __forceinline__ void apply_swap(cd2 (&a)[1 << R], int pair) {
    if constexpr (R == 4) {
        switch (pair) {
            case 0: QS_CASE("swap 01 <"); swap_bits<R, 0, 1>(a); QS_CASE("swap 01 >"); break;
            case 1: QS_CASE("swap 02 <"); swap_bits<R, 0, 2>(a); QS_CASE("swap 02 >"); break;
            case 2: QS_CASE("swap 03 <"); swap_bits<R, 0, 3>(a); QS_CASE("swap 03 >"); break;
            case 3: QS_CASE("swap 12 <"); swap_bits<R, 1, 2>(a); QS_CASE("swap 12 >"); break;
            case 4: QS_CASE("swap 13 <"); swap_bits<R, 1, 3>(a); QS_CASE("swap 13 >"); break;
            default: QS_CASE("swap 23 <"); swap_bits<R, 2, 3>(a); QS_CASE("swap 23 >"); break;
        }
    }
}

// General diagonal (non-unit entries or several controls), per amplitude; amplitudes
// without every control bit are multiplied by exactly 1.
template <int R>
__device__ __forceinline__ void apply_diag(cd2 (&a)[1 << R], int kind, int qa, int qb,
                                           uint64_t cm, const double* m, const Layout<R>& L,
                                           uint64_t base) {
    const cd2 one = mk2(1.0, 0.0);
    const bool two = kind == DG_DIAG2;
    // DG_DIAG1: entries d0, d1 by bit(qa); DG_DIAG2: d[2*bit(qa) + bit(qb)]
    const cd2 d0 = mk2(m[0], m[1]), d1 = mk2(m[2], m[3]);
    const cd2 d2 = mk2(m[4], m[5]), d3 = mk2(m[6], m[7]);
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        const uint64_t g = base | L.glob(j);
        const bool ok = (g & cm) == cm;
        const bool ba = (g >> qa) & 1ull;
        const bool bb = two && ((g >> qb) & 1ull);
        const cd2 d = two ? (ba ? (bb ? d3 : d2) : (bb ? d1 : d0)) : (ba ? d1 : d0);
        a[j] = cmul(ok ? d : one, a[j]);
    }
}

// Mask of register indices j (bit j set) whose register bit r is 1.
__device__ __forceinline__ constexpr unsigned reg_col_mask(int r) {
    return r == 0 ? 0xAAAAu : r == 1 ? 0xCCCCu : r == 2 ? 0xF0F0u : 0xFF00u;
}

// A unitary diagonal group (DevPhase).  The host folds the constant, thread-class,
// register and register-pair factors into one table entry per thread; block-class
// factors come from the prologue's per-tile table bf = [A_B, lambda_q, kappa_r]
// (lambda_q applies when thread bit q is set, kappa_r to register bit r).

// DG_PHASE_SIGN: every factor is +-1; the group is a 16-bit mask of the register
// indices to negate.  Thread t's mask is rebuilt from the group's descriptor (uniform,
// scalar loads) rather than read from a per-thread table: a vector load here would
// queue behind the tile's HBM traffic.
__device__ __forceinline__ unsigned sign_mask(const SignDesc& Q, int t) {
    unsigned M = Q.m0;
    // Uniform branches: most groups have no thread-bit or no thread-pair terms.
    unsigned any_mb = 0, any_pp = 0;
#pragma unroll
    for (int b = 0; b < SV_THREAD_BITS; ++b) {
        any_mb |= Q.mb[b];
        any_pp |= Q.pp[b];
    }
    if (any_mb) {
        // M ^ (all-ones if thread bit b is set) & mb[b]: one three-input bit op per bit
#pragma unroll
        for (int b = 0; b < SV_THREAD_BITS; ++b) M ^= (0u - ((unsigned)(t >> b) & 1u)) & Q.mb[b];
    }
    if (any_pp) {
        unsigned par = 0;
#pragma unroll
        for (int a = 0; a < SV_THREAD_BITS - 1; ++a)
            par ^= (unsigned)(t >> a) & (unsigned)__builtin_popcount(Q.pp[a] & (unsigned)t);
        M ^= 0u - (par & 1u);
    }
    return M & 0xFFFFu;
}

// DG_SIGN_REG: the negated register indices are the same in every thread, so the
// factors +-1.0 are formed on the scalar unit and each amplitude costs two FP64
// products with a scalar operand.
template <int R>
__device__ __forceinline__ void apply_sign_reg(cd2 (&a)[1 << R], unsigned M) {
    const unsigned one_hi = 0x3FF00000u;
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        const double s = __hiloint2double((int)(one_hi | (((M >> j) & 1u) << 31)), 0);
        a[j] = a[j] * s;
    }
}

template <int R>
__device__ __forceinline__ void apply_phase_sign(cd2 (&a)[1 << R], unsigned M,
                                                 const_cd2_ptr bf, int t) {
    if (bf) {
        // value 0 of a sign group: A_B and the packed negative lambda / kappa bits
        const cd2 b0 = bf[0];
        const unsigned bits = (unsigned)b0.y;
        constexpr unsigned kThreadMask = (1u << SV_THREAD_BITS) - 1u;
        const bool neg =
            (b0.x < 0.0) != (__builtin_popcount((unsigned)t & bits & kThreadMask) & 1);
        M ^= neg ? 0xFFFFu : 0u;
        if (bits & (1u << SV_THREAD_BITS)) M ^= reg_col_mask(0);
        if (R > 1 && (bits & (2u << SV_THREAD_BITS))) M ^= reg_col_mask(1);
        if (R > 2 && (bits & (4u << SV_THREAD_BITS))) M ^= reg_col_mask(2);
        if (R > 3 && (bits & (8u << SV_THREAD_BITS))) M ^= reg_col_mask(3);
    }
    // Factor +-1 of amplitude j: bit j of M shifted to bit 31 and XOR-ed into the sign
    // bits of the real and imaginary high words (one shift and two three-input bit ops,
    // v_bitop3, per amplitude; exact negation, as a product with -1.0 is).  Products
    // with +-1.0 took five instructions per amplitude: the factor's high word, a move of
    // its zero low word into a register pair, and two FP64 products.
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        const int s = (int)((M << (31 - j)) & 0x80000000u);
        a[j].x = __hiloint2double(__double2hiint(a[j].x) ^ s, __double2loint(a[j].x));
        a[j].y = __hiloint2double(__double2hiint(a[j].y) ^ s, __double2loint(a[j].y));
    }
}

// DG_PHASE: general unitary diagonal group.  The factor of register index j is
// A * prod_r kappa_r^{j_r} * prod_{r<s} c_rs^{j_r j_s}; it is built top bit by top bit
// (g[j + 2^t] = g[j] * kappa_t * prod_{r in j} c_rt) and applied once per amplitude.
// `e` is this thread's table entry [A_t, kappa_0..3].
template <int R>
__device__ __forceinline__ void apply_phase_cplx(cd2 (&a)[1 << R], const double* crr,
                                                 const cd2 (&e)[PHASE_THREAD_VALUES],
                                                 const_cd2_ptr bf, int t, int parts) {
    cd2 A = e[0], k0 = e[1], k1 = e[2], k2 = e[3], k3 = e[4];
    if (bf) {
        A = cmul(A, bf[0]);
        // Uniform branches: a group whose block terms all touch one register qubit (the
        // QFT's controlled phases after each H) has no thread x block factors.
        if (parts & 1) {
#pragma unroll
            for (int q = 0; q < SV_THREAD_BITS; ++q) {
                const cd2 Af = cmul(A, bf[1 + q]);
                A = ((t >> q) & 1) ? Af : A;
            }
        }
        if (parts & 2) {
            k0 = cmul(k0, bf[PHASE_KAPPA0]);
            if (R > 1) k1 = cmul(k1, bf[PHASE_KAPPA0 + 1]);
            if (R > 2) k2 = cmul(k2, bf[PHASE_KAPPA0 + 2]);
            if (R > 3) k3 = cmul(k3, bf[PHASE_KAPPA0 + 3]);
        }
    }
    // register-pair factors (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)
    const cd2 c01 = mk2(crr[0], crr[1]), c02 = mk2(crr[2], crr[3]);
    const cd2 c03 = mk2(crr[4], crr[5]), c12 = mk2(crr[6], crr[7]);
    const cd2 c13 = mk2(crr[8], crr[9]), c23 = mk2(crr[10], crr[11]);
    // Factors of indices 0..7 are kept for the top bit; each is applied once known.
    const cd2 g0 = A, g1 = cmul(A, k0);
    a[0] = cmul(g0, a[0]);
    a[1] = cmul(g1, a[1]);
    if constexpr (R > 1) {
        const cd2 g2 = cmul(A, k1), g3 = cmul(g1, cmul(k1, c01));
        a[2] = cmul(g2, a[2]);
        a[3] = cmul(g3, a[3]);
        if constexpr (R > 2) {
            __builtin_amdgcn_sched_barrier(0);
            const cd2 w1 = cmul(k2, c02), w2 = cmul(k2, c12), w3 = cmul(w1, c12);
            const cd2 g4 = cmul(A, k2), g5 = cmul(g1, w1);
            const cd2 g6 = cmul(g2, w2), g7 = cmul(g3, w3);
            a[4] = cmul(g4, a[4]);
            a[5] = cmul(g5, a[5]);
            a[6] = cmul(g6, a[6]);
            a[7] = cmul(g7, a[7]);
            if constexpr (R > 3) {
                __builtin_amdgcn_sched_barrier(0);
                const cd2 v1 = cmul(k3, c03), v2 = cmul(k3, c13), v4 = cmul(k3, c23);
                const cd2 v3 = cmul(v1, c13), v5 = cmul(v1, c23), v6 = cmul(v2, c23);
                const cd2 v7 = cmul(v3, c23);
                a[8] = cmul(cmul(A, k3), a[8]);
                a[9] = cmul(cmul(g1, v1), a[9]);
                a[10] = cmul(cmul(g2, v2), a[10]);
                a[11] = cmul(cmul(g3, v3), a[11]);
                __builtin_amdgcn_sched_barrier(0);
                a[12] = cmul(cmul(g4, v4), a[12]);
                a[13] = cmul(cmul(g5, v5), a[13]);
                a[14] = cmul(cmul(g6, v6), a[14]);
                a[15] = cmul(cmul(g7, v7), a[15]);
            }
        }
    }
}

// Product a * b where a (OA) or b (OB) is known, when the kernel is compiled, to be
// exactly 1: the other operand, which equals cmul(a, b) up to the sign of a zero part.
template <bool OA, bool OB>
__device__ __forceinline__ cd2 mul1(cd2 a, cd2 b) {
    if constexpr (OA) return b;
    else if constexpr (OB) return a;
    else return cmul(a, b);
}

// apply_phase_cplx for a per-plan kernel that knows which factors are exactly 1 (bit 0:
// A after its block factors, bits 1..4: kappa_0..3 after theirs, bits 5..10: the register
// pair factors c01 c02 c03 c12 c13 c23).  The same products in the same order with the
// products by 1 left out: a QFT group (controlled phases after one H) has one thread
// column and one register qubit's pair factors that are not 1.
template <int R, unsigned O>
__device__ __forceinline__ void apply_phase_cplx_known(cd2 (&a)[1 << R], const double* crr,
                                                       const cd2 (&e)[PHASE_THREAD_VALUES],
                                                       const_cd2_ptr bf, int t, int parts) {
    cd2 A = e[0], k0 = e[1], k1 = e[2], k2 = e[3], k3 = e[4];
    if (bf) {
        A = cmul(A, bf[0]);
        if (parts & 1) {
#pragma unroll
            for (int q = 0; q < SV_THREAD_BITS; ++q) {
                const cd2 Af = cmul(A, bf[1 + q]);
                A = ((t >> q) & 1) ? Af : A;
            }
        }
        if (parts & 2) {
            k0 = cmul(k0, bf[PHASE_KAPPA0]);
            if (R > 1) k1 = cmul(k1, bf[PHASE_KAPPA0 + 1]);
            if (R > 2) k2 = cmul(k2, bf[PHASE_KAPPA0 + 2]);
            if (R > 3) k3 = cmul(k3, bf[PHASE_KAPPA0 + 3]);
        }
    }
    constexpr bool oA = O & 1, o0 = (O >> 1) & 1, o1 = (O >> 2) & 1, o2 = (O >> 3) & 1,
                   o3 = (O >> 4) & 1;
    constexpr bool p01 = (O >> 5) & 1, p02 = (O >> 6) & 1, p03 = (O >> 7) & 1,
                   p12 = (O >> 8) & 1, p13 = (O >> 9) & 1, p23 = (O >> 10) & 1;
    const cd2 c01 = mk2(crr[0], crr[1]), c02 = mk2(crr[2], crr[3]);
    const cd2 c03 = mk2(crr[4], crr[5]), c12 = mk2(crr[6], crr[7]);
    const cd2 c13 = mk2(crr[8], crr[9]), c23 = mk2(crr[10], crr[11]);
    constexpr bool og1 = oA && o0;
    const cd2 g0 = A, g1 = mul1<oA, o0>(A, k0);
    if constexpr (!oA) a[0] = cmul(g0, a[0]);
    if constexpr (!og1) a[1] = cmul(g1, a[1]);
    if constexpr (R > 1) {
        constexpr bool ok1c = o1 && p01, og2 = oA && o1, og3 = og1 && ok1c;
        const cd2 g2 = mul1<oA, o1>(A, k1), g3 = mul1<og1, ok1c>(g1, mul1<o1, p01>(k1, c01));
        if constexpr (!og2) a[2] = cmul(g2, a[2]);
        if constexpr (!og3) a[3] = cmul(g3, a[3]);
        if constexpr (R > 2) {
            __builtin_amdgcn_sched_barrier(0);
            constexpr bool ow1 = o2 && p02, ow2 = o2 && p12, ow3 = ow1 && p12;
            const cd2 w1 = mul1<o2, p02>(k2, c02), w2 = mul1<o2, p12>(k2, c12);
            const cd2 w3 = mul1<ow1, p12>(w1, c12);
            constexpr bool og4 = oA && o2, og5 = og1 && ow1, og6 = og2 && ow2, og7 = og3 && ow3;
            const cd2 g4 = mul1<oA, o2>(A, k2), g5 = mul1<og1, ow1>(g1, w1);
            const cd2 g6 = mul1<og2, ow2>(g2, w2), g7 = mul1<og3, ow3>(g3, w3);
            if constexpr (!og4) a[4] = cmul(g4, a[4]);
            if constexpr (!og5) a[5] = cmul(g5, a[5]);
            if constexpr (!og6) a[6] = cmul(g6, a[6]);
            if constexpr (!og7) a[7] = cmul(g7, a[7]);
            if constexpr (R > 3) {
                __builtin_amdgcn_sched_barrier(0);
                constexpr bool ov1 = o3 && p03, ov2 = o3 && p13, ov4 = o3 && p23;
                constexpr bool ov3 = ov1 && p13, ov5 = ov1 && p23, ov6 = ov2 && p23,
                               ov7 = ov3 && p23;
                const cd2 v1 = mul1<o3, p03>(k3, c03), v2 = mul1<o3, p13>(k3, c13);
                const cd2 v4 = mul1<o3, p23>(k3, c23);
                const cd2 v3 = mul1<ov1, p13>(v1, c13), v5 = mul1<ov1, p23>(v1, c23);
                const cd2 v6 = mul1<ov2, p23>(v2, c23), v7 = mul1<ov3, p23>(v3, c23);
                if constexpr (!(oA && o3)) a[8] = cmul(mul1<oA, o3>(A, k3), a[8]);
                if constexpr (!(og1 && ov1)) a[9] = cmul(mul1<og1, ov1>(g1, v1), a[9]);
                if constexpr (!(og2 && ov2)) a[10] = cmul(mul1<og2, ov2>(g2, v2), a[10]);
                if constexpr (!(og3 && ov3)) a[11] = cmul(mul1<og3, ov3>(g3, v3), a[11]);
                __builtin_amdgcn_sched_barrier(0);
                if constexpr (!(og4 && ov4)) a[12] = cmul(mul1<og4, ov4>(g4, v4), a[12]);
                if constexpr (!(og5 && ov5)) a[13] = cmul(mul1<og5, ov5>(g5, v5), a[13]);
                if constexpr (!(og6 && ov6)) a[14] = cmul(mul1<og6, ov6>(g6, v6), a[14]);
                if constexpr (!(og7 && ov7)) a[15] = cmul(mul1<og7, ov7>(g7, v7), a[15]);
            }
        }
    }
}

// LDS barrier for the tile transposes: waits for this wave's LDS accesses, then syncs
// the workgroup.  The fences are workgroup-scope and restricted to the local (LDS)
// address space, so the barrier never waits for global memory operations in flight
// (unlike __syncthreads()), and the compiler keeps every LDS access on its side.  An
// inline-asm "s_waitcnt; s_barrier" with a memory clobber is not enough for the latter:
// with constant transpose offsets the compiler has moved a tile's LDS stores past such a
// barrier.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}


// DG_LAYOUT with both layouts known when the kernel is compiled (per-plan kernels): the
// register part of every local index is a constant and the thread part does not depend
// on the register index, and since the two parts occupy disjoint bits,
// lds_slot(thread | reg) = lds_slot(thread) + lds_slot(reg) (the low four bits cannot
// carry).  Each LDS access is then the thread's base address plus a constant offset (an
// immediate of the instruction), the same slots as op_body's DG_LAYOUT.
template <int R>
__device__ __forceinline__ void layout_const(cd2 (&a)[1 << R], Layout<R>& Lcur,
                                             const DevLayout& from, const DevLayout& to,
                                             cd2* __restrict__ lds, int t) {
    Layout<R> LA;
    make_layout<R>(LA, from, t);
    cd2* pa = lds + lds_slot(LA.loc_t);
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        int rl = 0;
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (j & (1 << i)) rl |= 1 << from.regs[i];
        pa[lds_slot(rl)] = a[j];
    }
    lds_barrier();
    make_layout<R>(Lcur, to, t);
    const cd2* pb = lds + lds_slot(Lcur.loc_t);
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        int rl = 0;
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (j & (1 << i)) rl |= 1 << to.regs[i];
        a[j] = pb[lds_slot(rl)];
    }
    lds_barrier();
}

// LDS of a per-plan kernel that transposes half a tile at a time (layout_split): 2048
// amplitudes plus one padding slot per 16.
constexpr int kLdsHalfSlots = kLdsSlots / 2;

// Local index l without bit p (the bits above p move down by one).
__device__ __forceinline__ constexpr int drop_bit(int l, int p) {
    return (l & ((1 << p) - 1)) | ((l >> (p + 1)) << p);
}

// Local index bits of register j under layout D.
template <int R>
__device__ __forceinline__ constexpr int reg_local(const DevLayout& D, int j) {
    int rl = 0;
#pragma unroll
    for (int i = 0; i < R; ++i)
        if (j & (1 << i)) rl |= 1 << D.regs[i];
    return rl;
}

// DG_LAYOUT between two compile-time layouts through half a tile of LDS: round h moves the
// amplitudes whose local bit p equals h.  Under a layout where p is register bit i, a
// thread writes (or reads) its 8 registers with bit i = h; where p is thread bit b, the
// threads with bit b of t = h write (or read) all 16.  Slots are indexed by the local
// index without bit p, and as in layout_const each access is the thread's base plus a
// constant (dropping bit p and the padding both keep disjoint bit sets additive).  With
// half the LDS of layout_const, up to four workgroups fit on a CU instead of two.
template <int R>
__device__ __forceinline__ void layout_split(cd2 (&a)[1 << R], Layout<R>& Lcur,
                                             const DevLayout& from, const DevLayout& to, int p,
                                             cd2* __restrict__ lds, int t) {
    constexpr int NR = 1 << R;
    int fi = -1, ti = -1, fb = -1, tb = -1;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        if (from.regs[i] == p) fi = i;
        if (to.regs[i] == p) ti = i;
    }
#pragma unroll
    for (int b = 0; b < SV_THREAD_BITS; ++b) {
        if (from.tpos[b] == p) fb = b;
        if (to.tpos[b] == p) tb = b;
    }
    Layout<R> LA;
    make_layout<R>(LA, from, t);
    make_layout<R>(Lcur, to, t);
    cd2* pa = lds + lds_slot(drop_bit(LA.loc_t & ~(1 << p), p));
    const cd2* pb = lds + lds_slot(drop_bit(Lcur.loc_t & ~(1 << p), p));
    cd2 b[NR];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (fi >= 0) {
#pragma unroll
            for (int j = 0; j < NR; ++j)
                if (((j >> fi) & 1) == h)
                    pa[lds_slot(drop_bit(reg_local<R>(from, j) & ~(1 << p), p))] = a[j];
        } else if (((t >> fb) & 1) == h) {
#pragma unroll
            for (int j = 0; j < NR; ++j) pa[lds_slot(drop_bit(reg_local<R>(from, j), p))] = a[j];
        }
        lds_barrier();
        if (ti >= 0) {
#pragma unroll
            for (int j = 0; j < NR; ++j)
                if (((j >> ti) & 1) == h)
                    b[j] = pb[lds_slot(drop_bit(reg_local<R>(to, j) & ~(1 << p), p))];
        } else if (((t >> tb) & 1) == h) {
#pragma unroll
            for (int j = 0; j < NR; ++j) b[j] = pb[lds_slot(drop_bit(reg_local<R>(to, j), p))];
        }
        lds_barrier();
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) a[j] = b[j];
}

// Runs one op of a pass on the register block.  H holds the record's head (kind through
// m[11]) and m the record's full matrix area (DevGate::m); `kind` equals H.kind.  The
// op-stream interpreter passes the head it loaded; a per-plan kernel passes the kind,
// register bits, controls and aux word as constants, so only the body of that op is
// compiled into it.  `lds` is the tile's transpose buffer (kLdsSlots slots), `base` the
// tile's global index bits.
template <int R>
__device__ __forceinline__ void op_body(cd2 (&a)[1 << R], Layout<R>& Lcur, const OpHead& H,
                                        const double* __restrict__ m, int kind,
                                        const cd2* __restrict__ tfactors,
                                        const cd2* __restrict__ tile_factors, uint32_t ntiles,
                                        uint64_t base, cd2* __restrict__ lds, int t) {
    constexpr int NR = 1 << R;
    const int r0 = H.r0, r1 = H.r1;
    if (kind >= DG_ROTS) {
        apply_rots<R>(a, kind, H.m, m);
        if (H.aux) apply_sign_reg<R>(a, (unsigned)H.aux);
    } else if (kind >= DG_ROT) {
        apply_rot<R>(a, kind, H.m[0], mk2(H.m[2], H.m[3]));
        if (H.aux) apply_sign_reg<R>(a, (unsigned)H.aux);
    } else if (kind == DG_SIGN_REG) {
        apply_sign_reg<R>(a, (unsigned)r0);
    } else if (kind == DG_U1X) {
        u1_bit<R, 0>(a, m);
        if constexpr (R > 1) u1_bit<R, (R > 1 ? 1 : 0)>(a, m + 8);
        if constexpr (R > 2) u1_bit<R, (R > 2 ? 2 : 0)>(a, m + 16);
        if constexpr (R > 3) u1_bit<R, (R > 3 ? 3 : 0)>(a, m + 24);
    } else if (kind == DG_U1S) {
        apply_u1s<R>(a, r0, m);
    } else if (kind == DG_PHASE_SIGN) {
        // the descriptor's 32-bit words from the head (word 2i is the low half of m[i])
        SignDesc Q;
        auto word = [&](int w) {
            return (unsigned)((w & 1) ? __double2hiint(H.m[w >> 1])
                                      : __double2loint(H.m[w >> 1]));
        };
        Q.m0 = word(0);
#pragma unroll
        for (int b = 0; b < SV_THREAD_BITS; ++b) {
            Q.mb[b] = word(1 + b);
            Q.pp[b] = word(1 + SV_THREAD_BITS + b);
        }
        Q.slot = (int)word(1 + 2 * SV_THREAD_BITS);
        const const_cd2_ptr bf =
            Q.slot >= 0 ? (const_cd2_ptr)(tile_factors +
                                          (size_t)Q.slot * ntiles * PHASE_BLOCK_VALUES)
                        : nullptr;
        apply_phase_sign<R>(a, sign_mask(Q, t), bf, t);
    } else if (kind == DG_PHASE) {
        const int tab = H.qb, slot = H.aux;
        const const_cd2_ptr bf =
            slot >= 0 ? (const_cd2_ptr)(tile_factors +
                                        (size_t)slot * ntiles * PHASE_BLOCK_VALUES)
                      : nullptr;
        const cd2* e = tfactors + (size_t)(tab + t) * PHASE_THREAD_VALUES;
        const cd2 f[PHASE_THREAD_VALUES] = {e[0], e[1], e[2], e[3], e[4]};
        apply_phase_cplx<R>(a, H.m, f, bf, t, r0);
    } else if (kind == DG_LAYOUT) {
#pragma unroll
        for (int j = 0; j < NR; ++j) lds[lds_slot(Lcur.loc(j))] = a[j];
        lds_barrier();
        make_layout<R>(Lcur, *reinterpret_cast<const DevLayout*>(m), t);
#pragma unroll
        for (int j = 0; j < NR; ++j) a[j] = lds[lds_slot(Lcur.loc(j))];
        lds_barrier();
    } else if (kind == DG_U1) {
        apply_cu1<R>(a, r0, H.cmask, m, Lcur, base);
    } else if (kind == DG_U2) {
        apply_u2<R>(a, r0, r1, H.cmask, m, Lcur, base);
    } else if (kind == DG_SWAP) {
        apply_swap<R>(a, r0);
    } else if (kind == DG_SCALE) {
        const cd2 f = mk2(H.m[0], H.m[1]);
        if (r0) {
#pragma unroll
            for (int j = 0; j < NR; ++j) a[j] = a[j] * f.x;
        } else {
#pragma unroll
            for (int j = 0; j < NR; ++j) a[j] = cmul(f, a[j]);
        }
    } else {
        apply_diag<R>(a, kind, H.qa, H.qb, H.cmask, m, Lcur, base);
    }
}

// Op head of a per-plan kernel's op: the fields the plan fixed are compile-time
// constants, the matrix words are read from the op record (m[0..11]).
__device__ __forceinline__ OpHead op_head(const DevGate* __restrict__ g, int kind, int r0,
                                          int r1, int qa, int qb, int aux, uint64_t cmask) {
    OpHead H;
    H.kind = kind;
    H.r0 = r0;
    H.r1 = r1;
    H.qa = qa;
    H.qb = qb;
    H.aux = aux;
    H.cmask = cmask;
#pragma unroll
    for (int i = 0; i < 12; ++i) H.m[i] = g->m[i];
    return H;
}

// Loads tile `base` of a pass into the register block under the pass's HBM load layout.
// The planner picks a load layout whose thread bits cover qubits 0..3 (16 lanes walk 256
// consecutive bytes).  The state streams through once per pass: nontemporal loads and
// stores keep it from displacing the op records and phase tables in the caches.
template <int R>
__device__ __forceinline__ void pass_load(cd2 (&a)[1 << R], Layout<R>& L,
                                          const cd2* __restrict__ state,
                                          const DevPass* __restrict__ P, uint64_t base, int t) {
    make_layout<R>(L, P->load, t);
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) a[j] = __builtin_nontemporal_load(state + (base | L.glob(j)));
}

// Stores the register block under the current layout (the op stream ends in a layout
// whose thread bits cover qubits 0..3).
template <int R>
__device__ __forceinline__ void pass_store(const cd2 (&a)[1 << R], const Layout<R>& L,
                                           cd2* __restrict__ state, uint64_t base) {
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) __builtin_nontemporal_store(a[j], state + (base | L.glob(j)));
}

}  // namespace qs
