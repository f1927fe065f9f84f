// Per-plan pass kernels compiled at run time (hipRTC, gfx950).
//
// The op-stream interpreter (sv_pass_kernel) reads every op's kind from the op table
// and branches to its body, once per op and workgroup.  For a plan that is launched many
// times, the dispatch can be compiled away: each pass gets a kernel that runs the pass's
// ops as one straight-line sequence of op_body calls whose kind, register bits, control
// mask and aux word are constants.  The op data (rotation parameters, phase
// descriptors, layouts) is still read from the uploaded op table, so two passes with the
// same op sequence share one kernel, and a plan of the same circuit structure with other
// angles reuses the compiled kernels of the process-wide cache.  The device code is
// sv_ops.h itself, embedded into the library at build time, so a specialised pass runs
// the interpreter's op bodies in the interpreter's order and its result is the same bit
// for bit.
#include <hip/hiprtc.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "qsim_internal.h"

namespace qs {

namespace {

// kSvDeviceTypesSrc, kSvOpsSrc: the text of sv_device_types.h and sv_ops.h (generated
// by the build from those files).
#include "sv_jit_src.inc"

// The fields of an op record that op_body reads as int fields of the head, by kind;
// the others are emitted as 0 so that equal op sequences give equal kernel text.
void head_constants(const DevGate& G, int* r0, int* r1, int* qa, int* qb, int* aux,
                    uint64_t* cmask) {
    *r0 = *r1 = *qa = *qb = *aux = 0;
    *cmask = 0;
    const int k = G.kind;
    if (k >= DG_ROT) {
        *aux = G.aux;
    } else if (k == DG_SIGN_REG || k == DG_U1S || k == DG_SWAP || k == DG_SCALE) {
        *r0 = G.r0;
    } else if (k == DG_PHASE) {
        *r0 = G.r0;
        *qb = G.qb;
        *aux = G.aux;
    } else if (k == DG_U1) {
        *r0 = G.r0;
        *cmask = G.cmask;
    } else if (k == DG_U2) {
        *r0 = G.r0;
        *r1 = G.r1;
        *cmask = G.cmask;
    } else if (k == DG_DIAG1 || k == DG_DIAG2) {
        *qa = G.qa;
        *qb = G.qb;
        *cmask = G.cmask;
    }
}

// A +-1 phase group (DG_PHASE_SIGN) with its sign descriptor as constants: the per-thread
// negation mask of sign_mask() reduces to the terms the group has (the interpreter
// rebuilds it from all eight thread-bit words and their pair rows), and its block factors
// are read from the group's slot.  The same mask as sign_mask(), so the same result.
std::string sign_group_code(const DevGate& G) {
    SignDesc Q;
    std::memcpy(&Q, G.m, sizeof(Q));
    char buf[160];
    std::string c = "    {\n";
    std::snprintf(buf, sizeof(buf), "        unsigned M = %uu;\n", Q.m0);
    c += buf;
    for (int b = 0; b < SV_THREAD_BITS; ++b) {
        if (!Q.mb[b]) continue;
        std::snprintf(buf, sizeof(buf),
                      "        M ^= (0u - ((unsigned)(t >> %d) & 1u)) & %uu;\n", b, Q.mb[b]);
        c += buf;
    }
    bool any_pp = false;
    for (int b = 0; b < SV_THREAD_BITS; ++b) any_pp = any_pp || Q.pp[b];
    if (any_pp) {
        c += "        unsigned par = 0;\n";
        for (int a = 0; a < SV_THREAD_BITS - 1; ++a) {
            if (!Q.pp[a]) continue;
            std::snprintf(buf, sizeof(buf),
                          "        par ^= (unsigned)(t >> %d) & (unsigned)__builtin_popcount(%uu & (unsigned)t);\n",
                          a, Q.pp[a]);
            c += buf;
        }
        c += "        M ^= 0u - (par & 1u);\n";
    }
    if (Q.slot >= 0) {
        std::snprintf(buf, sizeof(buf),
                      "        const const_cd2_ptr bf = (const_cd2_ptr)(tile_factors + (size_t)%d * ntiles * PHASE_BLOCK_VALUES);\n",
                      Q.slot);
        c += buf;
    } else {
        c += "        const const_cd2_ptr bf = nullptr;\n";
    }
    c += "        apply_phase_sign<R>(a, M & 0xFFFFu, bf, t);\n    }\n";
    return c;
}

// The launch bounds every pass kernel text carries.
constexpr const char* kLaunchBounds = "__launch_bounds__(1 << (qs::SV_TILE_BITS - qs::SV_REG_BITS))";

// A half-tile-LDS kernel (34 KiB, four workgroups per CU by LDS) that needs more than 168
// VGPRs runs two waves per SIMD; asking the compiler for three often fits it without
// spilling (the first light-cone pass: 172 -> 164 VGPRs, 3.9 -> 3.3 ms).
constexpr int kThreeWaveVgprs = 168;

// A DevLayout as a constant initializer (regs, rq, tpos, tq).
std::string layout_literal(const DevLayout& D) {
    std::string c = "{{";
    auto arr = [&](const int32_t* v, int n) {
        for (int i = 0; i < n; ++i) c += (i ? "," : "") + std::to_string(v[i]);
    };
    arr(D.regs, 4);
    c += "},{";
    arr(D.rq, 4);
    c += "},{";
    arr(D.tpos, SV_THREAD_BITS);
    c += "},{";
    arr(D.tq, SV_THREAD_BITS);
    c += "}}";
    return c;
}

// Local position along which layout_split halves the transpose from -> to (R register
// bits each): a position that is a register bit under both layouts keeps every thread
// busy in both rounds; otherwise a register bit of one layout that is a thread bit of the
// other, preferring the highest thread bit (bits 6 and 7 select the wave, so the threads
// that sit a round out are whole waves).
int split_position(const DevLayout& from, const DevLayout& to, int R) {
    for (int i = 0; i < R; ++i)
        for (int k = 0; k < R; ++k)
            if (from.regs[i] == to.regs[k]) return from.regs[i];
    int best = -1, best_bit = -1;
    for (int b = 0; b < SV_THREAD_BITS; ++b)
        for (int i = 0; i < R; ++i) {
            if (to.tpos[b] == from.regs[i] && b > best_bit) best = from.regs[i], best_bit = b;
            if (from.tpos[b] == to.regs[i] && b > best_bit) best = to.regs[i], best_bit = b;
        }
    return best;
}

// How entry e of thread 0's row of a complex phase group's table (column e mod
// PHASE_THREAD_VALUES) varies over the nthr threads: -1 when every thread has the same
// value (bit for bit), b when the value depends on thread bit b alone, SV_THREAD_BITS
// otherwise.
int thread_dependence(const SvPlan& plan, size_t e, int nthr) {
    const double* v = plan.thread_factors.data();
    auto at = [&](int t) { return v + 2 * (e + (size_t)t * PHASE_THREAD_VALUES); };
    auto same = [&](int t, int u) { return std::memcmp(at(t), at(u), 2 * sizeof(double)) == 0; };
    int bits = 0;
    for (int t = 1; t < nthr; ++t)
        if (!same(t, 0)) bits |= t;
    if (!bits) return -1;
    for (int b = 0; (1 << b) < nthr; ++b) {
        bool ok = true;
        for (int t = 0; t < nthr && ok; ++t) ok = same(t, t & (1 << b));
        if (ok) return b;
    }
    return SV_THREAD_BITS;
}

// Kernel text of pass P without its name (the name is spliced in at "@NAME@").
std::string pass_body(const SvPlan& plan, const DevPass& P) {
    std::string s;
    s.reserve(1024 + 240 * (size_t)P.nop);
    s += std::string("extern \"C\" __global__ void ") + kLaunchBounds + "\n";
    s += "@NAME@(qs::cd2* __restrict__ state, const qs::DevPass* __restrict__ P,\n"
         "       const qs::DevGate* __restrict__ gates, const qs::cd2* __restrict__ tfactors,\n"
         "       const qs::cd2* __restrict__ factors, uint32_t ntiles) {\n"
         "    using namespace qs;\n";
    s += "    constexpr int R = " + std::to_string(P.R) + ";\n";
    // Full tiles transpose half a tile at a time through LDS (layout_split).
    const bool split = P.K == SV_TILE_BITS && P.R == SV_REG_BITS;
    s += split ? "    __shared__ cd2 lds[kLdsHalfSlots];\n" : "    __shared__ cd2 lds[kLdsSlots];\n";
    s += "    const int t = threadIdx.x;\n"
         "    const uint32_t tile = blockIdx.x;\n"
         "    const cd2* tile_factors = factors + ((size_t)P->slot0 * ntiles + tile) * PHASE_BLOCK_VALUES;\n"
         "    const uint64_t base = tile_base(P, tile);\n"
         "    Layout<R> Lcur;\n"
         "    cd2 a[1 << R];\n"
         "    pass_load<R>(a, Lcur, state, P, base, t);\n"
         "    const DevGate* gt = gates + P->gate0;\n";
    char line[512];
    // Complex phase groups read 5 per-thread factors from a table (L2-resident): each
    // group's reads are issued after the previous group (the first group's after the tile
    // loads), so that they are in flight during the ops in between instead of costing
    // one round trip at the group.
    std::vector<int> phase_ops;
    for (int i = 0; i < P.nop; ++i)
        if (plan.gates[(size_t)P.gate0 + i].kind == DG_PHASE) phase_ops.push_back(i);
    // A table column (A_t or kappa_r(t)) that is the same for every thread, or depends on
    // one thread bit only, is read as one or two uniform values (scalar loads of thread 0's
    // and thread 2^b's entries) instead of one vector load per thread: the QFT's groups
    // have one column that depends on several thread bits and four that do not.  The
    // values are the table's own entries, so the group computes what the interpreter does.
    const int nthr = 1 << (P.K - P.R);
    auto table_loads = [&](int i) {
        const DevGate& G = plan.gates[(size_t)P.gate0 + i];
        std::snprintf(line, sizeof(line), "    cd2 ph%d[PHASE_THREAD_VALUES];\n", i);
        s += line;
        for (int k = 0; k < PHASE_THREAD_VALUES; ++k) {
            const size_t e = (size_t)G.qb * PHASE_THREAD_VALUES + k;
            switch (thread_dependence(plan, e, nthr)) {
                case -1:
                    std::snprintf(line, sizeof(line),
                                  "    ph%d[%d] = ((const_cd2_ptr)tfactors)[%zu];\n", i, k, e);
                    break;
                case SV_THREAD_BITS:
                    std::snprintf(line, sizeof(line),
                                  "    ph%d[%d] = tfactors[%zu + (size_t)t * PHASE_THREAD_VALUES];\n",
                                  i, k, e);
                    break;
                default: {
                    const int b = thread_dependence(plan, e, nthr);
                    std::snprintf(line, sizeof(line),
                                  "    ph%d[%d] = ((t >> %d) & 1) ? ((const_cd2_ptr)tfactors)[%zu] : ((const_cd2_ptr)tfactors)[%zu];\n",
                                  i, k, b, e + ((size_t)1 << b) * PHASE_THREAD_VALUES, e);
                }
            }
            s += line;
        }
    };
    // load point of each group's table: after op load_after[g], the previous group (-1:
    // before the first op).  Measured against loading at the group (0.3 ms per 28-qubit
    // step faster) and against also waiting for a transpose in between (as fast overall,
    // individual passes either way).
    std::vector<int> load_after(phase_ops.size());
    for (size_t g = 0; g < phase_ops.size(); ++g) load_after[g] = g ? phase_ops[g - 1] : -1;
    auto loads_after = [&](int i) {
        for (size_t g = 0; g < phase_ops.size(); ++g)
            if (load_after[g] == i) table_loads(phase_ops[g]);
    };
    loads_after(-1);
    // The register layout in force, known here: the pass's load layout, then each
    // DG_LAYOUT's target.  Transposes between two known layouts use constant LDS offsets.
    DevLayout cur_layout = P.load;
    for (int i = 0; i < P.nop; ++i) {
        const DevGate& G = plan.gates[(size_t)P.gate0 + i];
        if (i > 0) loads_after(i - 1);
        std::snprintf(line, sizeof(line), "    // op %d kind %d\n", i, G.kind);
        s += line;
        if (G.kind == DG_LAYOUT) {
            DevLayout to;
            std::memcpy(&to, G.m, sizeof(to));
            s += "    { constexpr DevLayout from = " + layout_literal(cur_layout) +
                 "; constexpr DevLayout to = " + layout_literal(to) + "; ";
            if (split) {
                std::snprintf(line, sizeof(line),
                              "layout_split<R>(a, Lcur, from, to, %d, lds, t); }\n",
                              split_position(cur_layout, to, P.R));
                s += line;
            } else {
                s += "layout_const<R>(a, Lcur, from, to, lds, t); }\n";
            }
            cur_layout = to;
            continue;
        }
        if (G.kind == DG_PHASE_SIGN) {
            s += sign_group_code(G);
            continue;
        }
        if (G.kind == DG_PHASE) {
            // op_body's DG_PHASE branch with the table entry loaded ahead and the factors
            // that are exactly 1 for every thread and tile left out of the products: a
            // table column that is the same 1 for every thread (unless block factors
            // multiply it), a register-pair factor of the record that is 1.
            unsigned ones = 0;
            for (int k = 0; k < PHASE_THREAD_VALUES; ++k) {
                const size_t e = (size_t)G.qb * PHASE_THREAD_VALUES + k;
                const double* v = plan.thread_factors.data() + 2 * e;
                if (thread_dependence(plan, e, nthr) == -1 && v[0] == 1.0 && v[1] == 0.0)
                    ones |= 1u << k;
            }
            if (G.aux >= 0) {
                ones &= ~1u;                              // A times the block factor A_B
                if (G.r0 & 2) ones &= 1u;                 // kappa_r times their block factors
            }
            for (int k = 0; k < 6; ++k)
                if (G.m[2 * k] == 1.0 && G.m[2 * k + 1] == 0.0) ones |= 1u << (5 + k);
            std::snprintf(line, sizeof(line),
                          "    { const OpHead H = op_head(gt + %d, %d, %d, 0, 0, %d, %d, 0ull); "
                          "const const_cd2_ptr bf = %s; "
                          "apply_phase_cplx_known<R, %uu>(a, H.m, ph%d, bf, t, %d); }\n",
                          i, G.kind, G.r0, G.qb, G.aux,
                          G.aux >= 0 ? "(const_cd2_ptr)(tile_factors + (size_t)H.aux * ntiles * PHASE_BLOCK_VALUES)"
                                     : "nullptr",
                          ones, i, G.r0);
            s += line;
            continue;
        }
        int r0, r1, qa, qb, aux;
        uint64_t cm;
        head_constants(G, &r0, &r1, &qa, &qb, &aux, &cm);
        if (G.kind >= DG_ROTS && aux != 0 && P.R == SV_REG_BITS) {
            // rotations, merged output factors and the register-only sign mask in one body
            std::snprintf(line, sizeof(line),
                          "    { const OpHead H = op_head(gt + %d, %d, 0, 0, 0, 0, %d, 0ull); "
                          "rots_body_signed<R, %d, %uu>(a, H.m, gt[%d].m + 4); }\n",
                          i, G.kind, aux, G.kind - DG_ROTS + 1, (unsigned)aux, i);
            s += line;
            continue;
        }
        std::snprintf(line, sizeof(line),
                      "    { const OpHead H = op_head(gt + %d, %d, %d, %d, %d, %d, %d, %lluull); "
                      "op_body<R>(a, Lcur, H, gt[%d].m, %d, tfactors, tile_factors, ntiles, base, lds, t); }\n",
                      i, G.kind, r0, r1, qa, qb, aux, (unsigned long long)cm, i, G.kind);
        s += line;
    }
    s += "    pass_store<R>(a, Lcur, state, base);\n}\n";
    return s;
}

std::string kernel_name(size_t i) { return "qs_sv_pass_" + std::to_string(i); }


// One hipRTC program over `bodies` (kernel i named kernel_name(i)); the code object into
// *code.  The log of a failed compile goes into the error message.
int compile_program(const std::vector<const std::string*>& bodies, std::vector<char>* code,
                    bool three_waves = false) {
    std::string src = "#include \"sv_ops.h\"\n";
    for (size_t i = 0; i < bodies.size(); ++i) {
        std::string b = *bodies[i];
        b.replace(b.find("@NAME@"), 6, kernel_name(i));
        if (three_waves) {
            const size_t at = b.find(kLaunchBounds);
            if (at != std::string::npos)
                b.insert(at + std::strlen(kLaunchBounds), " __attribute__((amdgpu_waves_per_eu(3)))");
        }
        src += b;
    }
    const char* headers[] = {kSvDeviceTypesSrc, kSvOpsSrc};
    const char* names[] = {"sv_device_types.h", "sv_ops.h"};
    hiprtcProgram prog = nullptr;
    hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "qs_sv_passes.hip", 2, headers, names);
    QS_REQUIRE(r == HIPRTC_SUCCESS, QS_E_HIP, "hiprtcCreateProgram failed: %s",
               hiprtcGetErrorString(r));
    // The same code generation as the library's own build: gfx950, -O3, HIP's default
    // contraction, no packed FP32 instructions.
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                          "-ffp-contract=fast-honor-pragmas", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops"};
    r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        std::string log;
        if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
            log.resize(n);
            (void)hiprtcGetProgramLog(prog, &log[0]);
        }
        (void)hiprtcDestroyProgram(&prog);
        set_error("hipRTC compile of %zu pass kernels failed: %s\n%.2000s", bodies.size(),
                  hiprtcGetErrorString(r), log.c_str());
        return QS_E_HIP;
    }
    size_t n = 0;
    r = hiprtcGetCodeSize(prog, &n);
    if (r == HIPRTC_SUCCESS) {
        code->resize(n);
        r = hiprtcGetCode(prog, code->data());
    }
    (void)hiprtcDestroyProgram(&prog);
    QS_REQUIRE(r == HIPRTC_SUCCESS, QS_E_HIP, "hipRTC code retrieval failed: %s",
               hiprtcGetErrorString(r));
    return QS_OK;
}

// Process-wide cache of loaded pass kernels: (device, kernel text) -> function.  Modules
// stay loaded until the process ends.
std::mutex g_cache_mu;
std::map<std::pair<int, std::string>, hipFunction_t> g_cache;

}  // namespace

std::string sv_jit_source(const SvPlan& plan, int pass) {
    std::string b = pass_body(plan, plan.passes[pass]);
    b.replace(b.find("@NAME@"), 6, kernel_name(0));
    return "#include \"sv_ops.h\"\n" + b;
}

int sv_jit_check(const SvPlan& plan, int* nkernels, size_t* code_bytes) {
    std::vector<std::string> uniq;
    for (const DevPass& P : plan.passes) {
        std::string b = pass_body(plan, P);
        if (std::find(uniq.begin(), uniq.end(), b) == uniq.end()) uniq.push_back(std::move(b));
    }
    std::vector<const std::string*> ptrs;
    for (const auto& b : uniq) ptrs.push_back(&b);
    size_t bytes = 0;
    if (!ptrs.empty()) {
        std::vector<char> code;
        const int st = compile_program(ptrs, &code);
        if (st) return st;
        bytes = code.size();
    }
    if (nkernels) *nkernels = (int)uniq.size();
    if (code_bytes) *code_bytes = bytes;
    return QS_OK;
}

int sv_jit_specialize(const SvPlan& plan, int device, std::vector<hipFunction_t>* fns,
                      int* compiled) {
    std::vector<std::string> body(plan.passes.size());
    for (size_t p = 0; p < plan.passes.size(); ++p) body[p] = pass_body(plan, plan.passes[p]);
    std::lock_guard<std::mutex> lock(g_cache_mu);
    std::vector<const std::string*> todo;
    for (const auto& b : body) {
        if (g_cache.count({device, b})) continue;
        if (std::find_if(todo.begin(), todo.end(), [&](const std::string* x) { return *x == b; }) ==
            todo.end())
            todo.push_back(&b);
    }
    if (!todo.empty()) {
        // One program for all of them: the run-time compiler serialises compilations
        // within a process, and one program parses the headers once.
        std::vector<char> code;
        const int st = compile_program(todo, &code);
        if (st) return st;
        hipModule_t mod = nullptr;
        QS_HIP(hipModuleLoadData(&mod, code.data()));
        std::vector<const std::string*> wide;
        for (size_t k = 0; k < todo.size(); ++k) {
            hipFunction_t f = nullptr;
            QS_HIP(hipModuleGetFunction(&f, mod, kernel_name(k).c_str()));
            g_cache[{device, *todo[k]}] = f;
            int vgprs = 0;
            QS_HIP(hipFuncGetAttribute(&vgprs, HIP_FUNC_ATTRIBUTE_NUM_REGS, f));
            if (vgprs > kThreeWaveVgprs && todo[k]->find("kLdsHalfSlots") != std::string::npos)
                wide.push_back(todo[k]);
        }
        if (!wide.empty()) {
            // Second compile of those kernels for three waves per SIMD; a kernel that then
            // spills keeps its first build.
            std::vector<char> code3;
            const int st3 = compile_program(wide, &code3, true);
            if (st3) return st3;
            hipModule_t mod3 = nullptr;
            QS_HIP(hipModuleLoadData(&mod3, code3.data()));
            for (size_t k = 0; k < wide.size(); ++k) {
                hipFunction_t f = nullptr;
                QS_HIP(hipModuleGetFunction(&f, mod3, kernel_name(k).c_str()));
                int scratch = 0, vgprs = 0;
                QS_HIP(hipFuncGetAttribute(&scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f));
                QS_HIP(hipFuncGetAttribute(&vgprs, HIP_FUNC_ATTRIBUTE_NUM_REGS, f));
                if (scratch == 0 && vgprs <= kThreeWaveVgprs) g_cache[{device, *wide[k]}] = f;
            }
        }
    }
    fns->resize(plan.passes.size());
    for (size_t p = 0; p < body.size(); ++p) (*fns)[p] = g_cache.at({device, body[p]});
    if (compiled) *compiled = (int)todo.size();
    return QS_OK;
}

}  // namespace qs
