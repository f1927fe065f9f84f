// Statevector data shared by the host planner, the pass kernels and the per-plan
// kernels compiled at run time: the complex type, the tile constants and the device
// records of a fused-pass plan.  No host library headers: the run-time compiler
// receives this file as it is.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#else
// the run-time compiler's own fixed-width integer types
using __hip_internal::int32_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#endif

namespace qs {

// Complex double as a native 2-vector.  HIP's double2 is a struct whose copies lower to
// memcpy, which keeps register arrays of it in scratch memory; this type stays in VGPRs.
typedef double cd2 __attribute__((ext_vector_type(2)));

__host__ __device__ __forceinline__ cd2 mk2(double re, double im) {
    cd2 r;
    r.x = re;
    r.y = im;
    return r;
}

constexpr int SV_MAX_QUBITS = 34;      // 2^34 * 16 B = 256 GiB, inside 288 GB of HBM
constexpr int SV_TILE_BITS = 12;       // qubits per fused pass tile (4096 amplitudes, 64 KiB)
constexpr int SV_REG_BITS = 4;         // amplitudes per thread = 16
constexpr int SV_LOW_BITS = 4;         // qubits 0..3 are in every tile (256 B runs)
constexpr int SV_THREAD_BITS = SV_TILE_BITS - SV_REG_BITS;   // thread-index bits of a full tile

// Device op of a fused pass (one record per op, executed in order by every thread).
//   DG_U1X    one-qubit gates on every register bit; the 2x2 matrix of register bit b
//             is m[8b .. 8b+7] (identity where r0, the mask of bits carrying a gate,
//             has no bit)
//   DG_U1S    uncontrolled one-qubit gate on register bit r0, matrix m[0..7]
//   DG_U1     controlled one-qubit gate on register bit r0, control mask cmask
//   DG_U2     two-qubit gate on register bits r0 (row-index MSB) and r1, controls cmask
//   DG_SWAP   uncontrolled SWAP of register-bit pair r0 of (0,1) (0,2) (0,3) (1,2)
//             (1,3) (2,3), four-bit register blocks only: a register permutation
//   DG_DIAG1/2 general diagonal on global qubits qa (MSB for DIAG2) and qb, evaluated
//             per amplitude (non-unit entries or several controls)
//   DG_PHASE / DG_PHASE_SIGN  unitary diagonal group, DevPhase record qa; the SIGN form
//             has only +-1 factors and carries its SignDesc at the start of m; DG_PHASE
//             carries what the kernel reads of the group: its per-thread table entry
//             qb, block-factor slot aux and register-pair factors crr in m[0..11];
//             r0 bit 0 / bit 1: the group has thread x block / register x block terms
//             (otherwise those block factors are ones and are not applied)
//   DG_SIGN_REG  a DG_PHASE_SIGN group whose factors depend on register bits only (no
//             thread-bit or block terms): amplitude j is negated when bit j of r0 is
//             set; m and qa as for DG_PHASE_SIGN
//   DG_LAYOUT transpose through LDS to the DevLayout stored at the start of m
//   DG_SCALE  every amplitude times (m[0], m[1]); r0 = 1 when the factor is real
//   DG_ROT + 4 r0 + r1 (16 kinds): scaled real rotation on register bit r0, t = m[0]:
//             r1 bit 0 clear maps a pair (x0, x1) to (x0 - t x1, x1 + t x0), set to
//             (t x0 - x1, x0 + t x1); r1 bit 1 set multiplies the second output by
//             k = (m[2], m[3]).  Bit and variant live in the kind so that the kernel
//             selects the straight-line body with one switch.  Nonzero aux (DG_ROT and
//             DG_ROTS): a DG_SIGN_REG mask applied after the rotations, with its phase
//             record in qa.
//   DG_ROTS + mask - 1 (15 kinds): first-form rotations with output factors on every
//             register bit of mask = r0 (1..15), one op: bit b has t = m[b] and output
//             factor k_b.  Rotations on distinct bits commute, and so do their output
//             factors, so the rotations run first and the factors follow as one
//             diagonal over the register index: amplitude j is multiplied by
//             f_j = (m[2 + 2j], m[3 + 2j]) (j = 1..15), the product of k_b over the bits b
//             of j & mask (one complex product per amplitude instead of one per
//             rotation and amplitude pair).
enum : int32_t {
    DG_U1 = 0, DG_U2 = 1, DG_DIAG1 = 2, DG_DIAG2 = 3, DG_U1X = 4, DG_PHASE = 5,
    DG_LAYOUT = 6, DG_PHASE_SIGN = 7, DG_U1S = 8, DG_SWAP = 9, DG_SCALE = 10, DG_SIGN_REG = 11,
    DG_ROT = 16,
    DG_ROTS = 32
};

// Matrix words of an op record: 32 for the largest matrices (DG_U2, DG_U1X), 34 for
// DG_ROTS (four rotation parameters and fifteen merged output factors).
constexpr int kDevGateWords = 36;

struct alignas(16) DevGate {
    int32_t kind;
    int32_t r0, r1;
    int32_t qa, qb;
    int32_t aux;           // DG_PHASE: the group's block-factor slot (-1: none)
    uint64_t cmask;        // global control mask
    double m[kDevGateWords];
};

// A register layout of a pass, precomputed on the host: which tile positions are the
// register bits, and where each thread-index bit lands (local position and global
// qubit).  Thread bits fill the non-register positions in ascending order; entries
// beyond the thread count are zero (those thread bits are always 0).
struct DevLayout {
    int32_t regs[4];       // local positions of register bits 0..R-1, ascending
    int32_t rq[4];         // global qubits of register bits
    int32_t tpos[SV_THREAD_BITS];   // local position of thread bit b
    int32_t tq[SV_THREAD_BITS];     // global qubit of thread bit b
};
static_assert(sizeof(DevLayout) <= sizeof(double) * 32, "DevLayout must fit in DevGate::m");

// A unitary diagonal group is a product of phase factors f^(bit products).  Factors are
// classified by where their qubits live in the pass: block bits (not in the tile),
// thread bits (tile positions that index the thread) or register bits.
//   block classes, folded per workgroup by sv_phase_block_kernel into
//   PHASE_BLOCK_VALUES factors [A_B, lambda_b (thread bits b), kappa_r (register bits r)]:
//     PT_BB: a, b global (b < 0: one-body)       -> A_B
//     PT_BT: a thread bit, b global              -> lambda_a
//     PT_BR: a register bit, b global            -> kappa_a
//   thread classes, evaluated per thread:
//     PT_T1: a thread bit                        -> A
//     PT_TT: a, b thread bits                    -> A
//     PT_TR: a register bit, b thread bit        -> kappa_a
enum : int32_t { PT_BB = 0, PT_BT = 1, PT_BR = 2, PT_T1 = 3, PT_TT = 4, PT_TR = 5 };
constexpr int PHASE_BLOCK_VALUES = 1 + SV_THREAD_BITS + SV_REG_BITS;
constexpr int PHASE_KAPPA0 = 1 + SV_THREAD_BITS;   // index of kappa_0
constexpr int PHASE_THREAD_VALUES = 5;

struct alignas(16) DevTerm {
    int32_t cls, a, b, pad;
    double re, im;
};

// The thread- and register-class part of a group is folded on the host into a table
// with one entry per thread: for SIGN groups a 16-bit mask of the register indices to
// negate (constant, thread, register and register-pair factors together), otherwise
// the complex factors (A_t, kappa_0..3 of thread t), PHASE_THREAD_VALUES per thread.
// The SIGN-group part of a DevPhase, also written at the start of the group's op
// record (DevGate::m) so that it arrives with the op head in one batch of scalar loads.
struct SignDesc {
    uint32_t m0;
    uint32_t mb[SV_THREAD_BITS];
    uint32_t pp[SV_THREAD_BITS];
    int32_t slot;
};

struct alignas(16) DevPhase {
    int32_t t0, nt;        // thread-class terms in the pass's term table
    int32_t b0, nb;        // block-class terms in the pass's term table
    int32_t slot;          // block-factor slot of this group in the pass (-1: none)
    int32_t sign;          // 1: every factor is +1 or -1 (parity arithmetic)
    int32_t tab;           // first entry of this group's per-thread table
    int32_t pad;
    // SIGN groups: the per-thread mask as the kernel rebuilds it, with t the thread
    // index: m0 ^ (XOR of mb[b] over set bits b of t) ^ (0xFFFF if the parity of
    // t_a t_b over the thread-pair terms is odd); bit b' of pp[a] marks the pair
    // (a, b'), a < b'.
    uint32_t m0;
    uint32_t mb[SV_THREAD_BITS];
    uint32_t pp[SV_THREAD_BITS];
    double c[2];           // constant factor
    double kr[4][2];       // constant per-register-bit factors
    double crr[6][2];      // register-pair factors: (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)
};

struct DevPass {
    int32_t K, R, nstage, tile_is_all;
    int32_t tq[SV_TILE_BITS];  // tile qubits ascending; local bit i <-> global qubit tq[i]
    uint64_t tile_mask;
    int32_t stage0;            // first stage in the stage array
    int32_t gate0;             // first op in the op array
    int32_t phase0;            // first phase record
    int32_t term0;             // first term
    int32_t nphase;            // phase records of this pass
    int32_t nslot;             // block-factor slots of this pass
    int32_t slot0;             // first slot in the block-factor table
    int32_t nop;               // ops of this pass
    int32_t nbq;               // block qubits (not in the tile): nq - K, or 0
    int32_t bq[SV_MAX_QUBITS]; // block qubits ascending: bit i of the tile index -> bq[i]
    // The same map as runs of consecutive block qubits: the tile's global index bits are
    // the OR over k of ((tile >> run_src[k]) & run_mask[k]) << run_dst[k]; unused runs
    // have mask 0.
    int32_t run_src[SV_TILE_BITS + 1];
    int32_t run_dst[SV_TILE_BITS + 1];
    uint32_t run_mask[SV_TILE_BITS + 1];
    DevLayout load;            // HBM load layout (the store layout is the last one set)
};

}  // namespace qs
