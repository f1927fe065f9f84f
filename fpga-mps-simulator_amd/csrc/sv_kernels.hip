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
#include "sv_ops.h"

namespace qs {

namespace {


// Block-class phase factors of every phase group of one pass, one thread per
// (slot, workgroup): [A_B, lambda_q (thread bits q), kappa_r (register bits r)].
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
    const uint64_t base = tile_base(P, blk);
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
            const int idx = (T.cls == PT_BT) ? 1 + T.a : PHASE_KAPPA0 + T.a;
#pragma unroll
            for (int i = 1; i < PHASE_BLOCK_VALUES; ++i)
                if (i == idx) v[i] = cmul(v[i], f);
        }
    }
    if (Q->sign) {
        // Sign groups read only value 0: its imaginary part (zero for a +-1 factor)
        // carries the negative lambda_q as bits q and negative kappa_r as bits
        // SV_THREAD_BITS + r.
        uint32_t bits = 0;
#pragma unroll
        for (int i = 1; i < PHASE_BLOCK_VALUES; ++i)
            if (v[i].x < 0.0) bits |= 1u << (i - 1);
        v[0].y = (double)bits;
    }
    cd2* out = factors + ((size_t)(P->slot0 + slot) * nblocks + blk) * PHASE_BLOCK_VALUES;
#pragma unroll
    for (int k = 0; k < PHASE_BLOCK_VALUES; ++k) out[k] = v[k];
}

// Runs a pass's op stream on one tile's register block: the op-stream interpreter.
// `lds` is the tile's transpose buffer (kLdsSlots slots); `base` the tile's global
// index bits.
template <int R>
__device__ __forceinline__ void run_ops(cd2 (&a)[1 << R], Layout<R>& Lcur,
                                        const DevGate* __restrict__ gt, int nop,
                                        const cd2* __restrict__ tfactors,
                                        const cd2* __restrict__ tile_factors, uint32_t ntiles,
                                        uint64_t base, cd2* __restrict__ lds, int t) {
    // Every op is straight-line code on the register block.  The dispatch is unrolled
    // twice: an op writes its results to fresh registers while its inputs are live, so
    // two ops per trip let the results ping-pong between two register sets instead of
    // being copied back at every loop latch.
    auto run_op = [&](int gi) __attribute__((always_inline)) {
        const DevGate& G = gt[gi];
        // The first 128 bytes of the record in one batch of scalar loads: the frequent
        // ops (DG_ROT, the phase groups, DG_SCALE) need nothing else, so they cost one
        // scalar-load round trip instead of one for the kind and one for the operands.
        const OpHead H = *reinterpret_cast<const OpHead*>(&G);
        // consumed here, so all of them are loaded (and waited for) before the branch
        asm volatile("" ::"s"(H.kind), "s"(H.r0), "s"(H.r1), "s"(H.qa), "s"(H.qb), "s"(H.aux),
                     "s"(H.m[0]), "s"(H.m[1]), "s"(H.m[2]), "s"(H.m[3]), "s"(H.m[4]),
                     "s"(H.m[5]), "s"(H.m[6]), "s"(H.m[7]), "s"(H.m[8]), "s"(H.m[9]),
                     "s"(H.m[10]), "s"(H.m[11]));
        op_body<R>(a, Lcur, H, G.m, H.kind, tfactors, tile_factors, ntiles, base, lds, t);
    };
    for (int gi = 0; gi < nop; gi += 2) {
        run_op(gi);
        if (gi + 1 < nop) run_op(gi + 1);
    }
}


// One fused pass, one tile per workgroup: workgroup `tile` loads its 2^K amplitudes,
// runs the pass's op stream on them in registers (LDS only for layout changes) and
// stores them back.
template <int R>
__global__ void __launch_bounds__(1 << (SV_TILE_BITS - SV_REG_BITS))
sv_pass_kernel(cd2* __restrict__ state, const DevPass* __restrict__ P,
               const DevGate* __restrict__ gates, const cd2* __restrict__ tfactors,
               const cd2* __restrict__ factors, uint32_t ntiles) {
    __shared__ cd2 lds[kLdsSlots];
    const int t = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    const cd2* tile_factors = factors + ((size_t)P->slot0 * ntiles + tile) * PHASE_BLOCK_VALUES;
    const uint64_t base = tile_base(P, tile);
    Layout<R> Lcur;
    cd2 a[1 << R];
    pass_load<R>(a, Lcur, state, P, base, t);
    run_ops<R>(a, Lcur, gates + P->gate0, P->nop, tfactors, tile_factors, ntiles, base, lds,
               t);
    pass_store<R>(a, Lcur, state, base);
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

// Pass-kernel timing (qs_sv_time_passes): events from a per-thread pool, bracketing
// each pass launch on its stream.
struct PassTimer {
    bool on = false;
    std::vector<hipEvent_t> pool;      // created events, reused
    size_t used = 0;                   // events recorded since the last read-out
    ~PassTimer() {
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
thread_local PassTimer g_timer;

int sv_plan_block_factors(int nq, const SvPlan& plan, const void* dev_blob, hipStream_t stream) {
    const char* blob = reinterpret_cast<const char*>(dev_blob);
    const SvBlobLayout L = sv_blob_layout(plan, nq);
    const DevPass* dpasses = reinterpret_cast<const DevPass*>(blob + L.passes);
    const DevPhase* dphases = reinterpret_cast<const DevPhase*>(blob + L.phases);
    const DevTerm* dterms = reinterpret_cast<const DevTerm*>(blob + L.terms);
    cd2* dfactors = reinterpret_cast<cd2*>(const_cast<char*>(blob) + L.factors);
    for (size_t p = 0; p < plan.passes.size(); ++p) {
        const DevPass& P = plan.passes[p];
        if (P.nslot <= 0) continue;
        const unsigned tiles = P.tile_is_all ? 1u : (1u << (nq - P.K));
        const uint64_t work = (uint64_t)P.nslot * tiles;
        hipLaunchKernelGGL(sv_phase_block_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256),
                           0, stream, nq, dpasses + p, dphases, dterms, dfactors, tiles);
        QS_HIP(hipGetLastError());
    }
    return QS_OK;
}

int sv_launch_plan(double2* state, int nq, const SvPlan& plan, const void* dev_blob,
                   hipStream_t stream, const hipFunction_t* spec, bool block_factors) {
    const char* blob = reinterpret_cast<const char*>(dev_blob);
    const SvBlobLayout L = sv_blob_layout(plan, nq);
    const DevPass* dpasses = reinterpret_cast<const DevPass*>(blob + L.passes);
    const DevGate* dgates = reinterpret_cast<const DevGate*>(blob + L.gates);
    const DevPhase* dphases = reinterpret_cast<const DevPhase*>(blob + L.phases);
    const DevTerm* dterms = reinterpret_cast<const DevTerm*>(blob + L.terms);
    const cd2* dtf = reinterpret_cast<const cd2*>(blob + L.tfactors);
    cd2* dfactors = reinterpret_cast<cd2*>(const_cast<char*>(blob) + L.factors);
    cd2* st = reinterpret_cast<cd2*>(state);
    for (size_t p = 0; p < plan.passes.size(); ++p) {
        const DevPass& P = plan.passes[p];
        const DevPass* dP = dpasses + p;
        const unsigned tiles = P.tile_is_all ? 1u : (1u << (nq - P.K));
        const unsigned threads = 1u << (P.K - P.R);
        const unsigned grid = tiles;
        if (block_factors && P.nslot > 0) {
            const uint64_t work = (uint64_t)P.nslot * tiles;
            hipLaunchKernelGGL(sv_phase_block_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256),
                               0, stream, nq, dP, dphases, dterms, dfactors, tiles);
            QS_HIP(hipGetLastError());
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (g_timer.on) {
            e0 = g_timer.next();
            e1 = g_timer.next();
            QS_REQUIRE(e0 && e1, QS_E_HIP, "hipEventCreate failed");
            QS_HIP(hipEventRecord(e0, stream));
        }
        if (spec && spec[p]) {
            // the pass's own kernel (sv_jit.cpp): the same arguments as sv_pass_kernel
            const DevPass* dPv = dP;
            const DevGate* dg = dgates;
            const cd2* tf = dtf;
            cd2* fa = dfactors;
            uint32_t nt = tiles;
            void* args[] = {&st, &dPv, &dg, &tf, &fa, &nt};
            QS_HIP(hipModuleLaunchKernel(spec[p], grid, 1, 1, threads, 1, 1, 0, stream, args,
                                         nullptr));
        } else switch (P.R) {
            case 1: hipLaunchKernelGGL(sv_pass_kernel<1>, dim3(grid), dim3(threads), 0, stream,
                                       st, dP, dgates, dtf, dfactors, tiles); break;
            case 2: hipLaunchKernelGGL(sv_pass_kernel<2>, dim3(grid), dim3(threads), 0, stream,
                                       st, dP, dgates, dtf, dfactors, tiles); break;
            case 3: hipLaunchKernelGGL(sv_pass_kernel<3>, dim3(grid), dim3(threads), 0, stream,
                                       st, dP, dgates, dtf, dfactors, tiles); break;
            default: hipLaunchKernelGGL(sv_pass_kernel<4>, dim3(grid), dim3(threads), 0, stream,
                                        st, dP, dgates, dtf, dfactors, tiles); break;
        }
        QS_HIP(hipGetLastError());
        if (e1) QS_HIP(hipEventRecord(e1, stream));
    }
    return QS_OK;
}

int sv_time_passes(int enable) {
    g_timer.on = enable != 0;
    g_timer.used = 0;
    return QS_OK;
}

int sv_pass_times(double* ms_out, int cap, int* count) {
    const int n = (int)(g_timer.used / 2);
    for (int i = 0; i < n && i < cap; ++i) {
        QS_HIP(hipEventSynchronize(g_timer.pool[2 * i + 1]));
        float ms = 0.0f;
        QS_HIP(hipEventElapsedTime(&ms, g_timer.pool[2 * i], g_timer.pool[2 * i + 1]));
        ms_out[i] = ms;
    }
    *count = n;
    g_timer.used = 0;
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
