// Internal helpers shared by the HIP translation units of libqsim_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/qsim.h"
#include "sv_device_types.h"

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

// Host-to-device copy of `bytes` from host memory that may be pageable, on `stream`.
// *done receives an event recorded after the copy (one per host thread, reused by the
// next call); h2d_copy_wait(*done) returns once the copy has read the source, so the
// caller keeps the source alive until then and may queue more work first.
int h2d_copy(void* dst, const void* src, size_t bytes, hipStream_t stream, hipEvent_t* done);
int h2d_copy_wait(hipEvent_t done);
// Host-to-device copy through a pinned staging ring of the calling thread (8 buffers, each
// reused once the copy that last read it has completed): `src` may be released as soon as
// the call returns, and the call does not wait for the stream.  Allocates pinned host
// memory (hipHostMalloc) when a buffer has to grow.
int h2d_copy_staged(void* dst, const void* src, size_t bytes, hipStream_t stream);

// Resolves the device address of the MPS static scratch array on `device`, which must be
// the current device (qs_init calls it); later calls for the same device do nothing.
int mps_scratch_setup(int device);

// ----------------------------------------------------------------- statevector

// Planner-side stage: register positions and the op range run under them (the op
// range starts with the DG_LAYOUT op that establishes the layout, except for a first
// stage loaded directly from HBM).  The device reads only the op stream.
struct DevStage {
    int32_t regs[4];       // local tile positions of the register bits, ascending
    int32_t g0, g1;        // op range in the pass's op array
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
    std::vector<DevPhase> phases;
    std::vector<DevTerm> terms;
    std::vector<double> thread_factors;   // complex groups: 2 * PHASE_THREAD_VALUES per thread
    std::vector<uint32_t> thread_masks;   // sign groups: one mask per thread (host copy
                                          // for plan inspection; the kernel rebuilds it)
    int nslot_total = 0;       // block-factor slots over all passes
};

// Device blob layout of an uploaded plan: [passes][stages][gates][phases][terms], each
// table starting on a 256-byte boundary, followed by the block-factor table
// (nslot_total x 2^(nq-K) x 13 complex) written by the prologue kernels.
inline size_t sv_blob_offset(size_t bytes) { return ((bytes + 255) / 256) * 256; }

struct SvBlobLayout {
    size_t passes, stages, gates, phases, terms, tfactors, factors, total;
};
SvBlobLayout sv_blob_layout(const SvPlan& plan, int nq);

// Builds the fused-pass plan.  Returns QS_OK or an error code (message set).
int sv_build_plan(int nq, const qs_gate* gates, int ng, SvPlan* plan);

// Launches every pass of `plan` (already uploaded to `dev_blob`) on `stream`.  spec, when
// not null, holds per pass a kernel compiled for that pass (sv_jit_specialize) or null
// for the op-stream interpreter.
// With block_factors, each pass's block-factor table is written by its prologue kernel
// first; a plan whose table was written once (sv_plan_block_factors) passes false.
int sv_launch_plan(double2* state, int nq, const SvPlan& plan, const void* dev_blob,
                   hipStream_t stream, const hipFunction_t* spec = nullptr,
                   bool block_factors = true);

// Writes the block-factor tables of every pass of an uploaded plan on `stream` (one
// launch per pass with block-class phase terms).  They depend on the plan only.
int sv_plan_block_factors(int nq, const SvPlan& plan, const void* dev_blob, hipStream_t stream);

// Per-plan pass kernels (sv_jit.cpp).  sv_jit_specialize compiles, with hipRTC for
// gfx950, the kernels of the plan's passes that the process-wide cache lacks, loads them
// on `device` (the current device) and writes one function per pass into *fns;
// *compiled receives the number of kernels compiled by this call.  sv_jit_check
// generates and compiles the plan's kernels without loading them (no device needed):
// *nkernels distinct kernels, *code_bytes of code objects.
int sv_jit_specialize(const SvPlan& plan, int device, std::vector<hipFunction_t>* fns,
                      int* compiled);
int sv_jit_check(const SvPlan& plan, int* nkernels, size_t* code_bytes);
// Source text of pass `pass`'s kernel (named qs_sv_pass_0), as sv_jit_specialize compiles it.
std::string sv_jit_source(const SvPlan& plan, int pass);

// Fills a DevLayout for register positions regs[0..R-1] (ascending) of pass P.
void sv_fill_layout(const DevPass& P, const int* regs, DevLayout* out);

// Pass-kernel timing behind qs_sv_time_passes / qs_sv_pass_times.
int sv_time_passes(int enable);
int sv_pass_times(double* ms_out, int cap, int* count);

int sv_fill_zero(double2* state, int nq, hipStream_t stream);
int sv_scale(double2* state, int nq, double re, double im, hipStream_t stream);
int sv_project(double2* state, int nq, int q, hipStream_t stream);
int sv_probs(const double2* state, int nq, double* p, hipStream_t stream);
int sv_norm_sq(const double2* state, int nq, double* out, hipStream_t stream);
// Measurement shots (sv_sample.hip): workspace size, and the sampler (synchronous for the
// normalisation check; the indices are written on `stream`).
size_t sv_sample_workspace_bytes(int nq);
int sv_sample(const double2* state, int nq, const double* uniforms, int shots, int64_t* out,
              double* ws, size_t ws_bytes, hipStream_t stream);

}  // namespace qs
