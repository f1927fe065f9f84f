// Internal helpers shared by the HIP translation units of libqsim_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/qsim.h"

namespace qs {

// Thread-local error message, returned by qs_last_error().
void set_error(const char* fmt, ...);
void clear_error();

#define QS_HIP(call)                                                              \
    do {                                                                          \
        hipError_t qs_e_ = (call);                                                \
        if (qs_e_ != hipSuccess) {                                                \
            ::qs::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(qs_e_), \
                            __FILE__, __LINE__);                                  \
            return QS_E_HIP;                                                      \
        }                                                                         \
    } while (0)

#define QS_REQUIRE(cond, code, ...)                                               \
    do {                                                                          \
        if (!(cond)) {                                                            \
            ::qs::set_error(__VA_ARGS__);                                         \
            return (code);                                                        \
        }                                                                         \
    } while (0)

// ----------------------------------------------------------------- statevector

constexpr int SV_MAX_QUBITS = 34;      // 2^34 * 16 B = 256 GiB, inside 288 GB of HBM
constexpr int SV_TILE_BITS = 12;       // qubits per fused pass tile (4096 amplitudes)
constexpr int SV_REG_BITS = 4;         // amplitudes per thread = 16
constexpr int SV_LOW_BITS = 4;         // qubits 0..3 are in every tile (256 B runs)
constexpr int SV_MAX_STAGES = 24;      // register layouts per pass

// Device gate of a fused pass.  U1/U2 name register bits (positions within the
// thread's 2^R register block); DIAG names global qubits.
enum : int32_t { DG_U1 = 0, DG_U2 = 1, DG_DIAG1 = 2, DG_DIAG2 = 3, DG_SIGN2 = 4 };

struct alignas(16) DevGate {
    int32_t kind;
    int32_t r0, r1;        // register bits (U1: r0; U2: r0 = MSB of row index)
    int32_t qa, qb;        // DIAG: global qubits (qa = MSB for DIAG2)
    int32_t pad;
    uint64_t cmask;        // global control mask
    double m[32];
};

struct DevStage {
    int32_t regs[4];       // local tile positions of the register bits, ascending
    int32_t g0, g1;        // gate range in the pass's gate array
};

struct DevPass {
    int32_t K, R, nstage, tile_is_all;
    int32_t tq[SV_TILE_BITS];  // tile qubits ascending; local bit i <-> global qubit tq[i]
    uint64_t tile_mask;
    int32_t stage0;            // first stage in the stage array
    int32_t gate0;             // first gate in the gate array
};

// Host-side logical gate, after name -> matrix conversion.
struct HostGate {
    int32_t kind;          // QS_GATE_*
    int32_t q0, q1;
    uint64_t cmask;
    double m[32];
};

struct SvPlan {
    std::vector<DevPass> passes;
    std::vector<DevStage> stages;
    std::vector<DevGate> gates;
};

// Device blob layout of an uploaded plan: [passes][stages][gates], each table
// starting on a 256-byte boundary.
inline size_t sv_blob_offset(size_t bytes) { return ((bytes + 255) / 256) * 256; }

// Builds the fused-pass plan.  Returns QS_OK or an error code (message set).
int sv_build_plan(int nq, const qs_gate* gates, int ng, SvPlan* plan);

// Launches every pass of `plan` (already uploaded to `dev_blob`) on `stream`.
int sv_launch_plan(double2* state, int nq, const SvPlan& plan, const void* dev_blob,
                   hipStream_t stream);

int sv_fill_zero(double2* state, int nq, hipStream_t stream);
int sv_scale(double2* state, int nq, double re, double im, hipStream_t stream);
int sv_project(double2* state, int nq, int q, hipStream_t stream);
int sv_probs(const double2* state, int nq, double* p, hipStream_t stream);
int sv_norm_sq(const double2* state, int nq, double* out, hipStream_t stream);

}  // namespace qs
