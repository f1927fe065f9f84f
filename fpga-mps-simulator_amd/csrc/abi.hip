// extern "C" entry points of libqsim_hip.so (declared in include/qsim.h).
#include <cmath>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>

#include "qsim_internal.h"

namespace qs {

namespace {
thread_local std::string g_last_error;
}

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

namespace {
// One event per host thread marks the completion of that thread's last upload.
struct CopyEvent {
    hipEvent_t e = nullptr;
    ~CopyEvent() {
        if (e) (void)hipEventDestroy(e);
    }
};
thread_local CopyEvent g_copy_event;
}  // namespace

int h2d_copy(void* dst, const void* src, size_t bytes, hipStream_t stream, hipEvent_t* done) {
    *done = nullptr;
    if (!g_copy_event.e) QS_HIP(hipEventCreateWithFlags(&g_copy_event.e, hipEventDisableTiming));
    QS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    QS_HIP(hipEventRecord(g_copy_event.e, stream));
    *done = g_copy_event.e;
    return QS_OK;
}

int h2d_copy_wait(hipEvent_t done) {
    QS_HIP(hipEventSynchronize(done));
    return QS_OK;
}

namespace {
// Pinned staging ring of one host thread: kRing buffers, each with the event recorded
// after the copy that last read it.  A buffer is reused only once that copy has completed,
// so the caller's host memory is free as soon as it has been copied into the ring.  The
// buffers are kept until the process ends (no destructor: at exit the HIP runtime may
// already be torn down, and the driver releases pinned memory with the process).
constexpr int kRing = 8;
struct StagingRing {
    void* buf[kRing] = {};
    size_t cap[kRing] = {};
    hipEvent_t ev[kRing] = {};
    bool used[kRing] = {};
    int next = 0;
};
thread_local StagingRing g_ring;
}  // namespace

int h2d_copy_staged(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    if (bytes == 0) return QS_OK;
    StagingRing& R = g_ring;
    const int k = R.next;
    R.next = (k + 1) % kRing;
    if (!R.ev[k]) QS_HIP(hipEventCreateWithFlags(&R.ev[k], hipEventDisableTiming));
    if (R.used[k]) QS_HIP(hipEventSynchronize(R.ev[k]));   // the copy that last read it
    if (R.cap[k] < bytes) {
        if (R.buf[k]) QS_HIP(hipHostFree(R.buf[k]));
        R.buf[k] = nullptr;
        R.cap[k] = 0;
        size_t cap = 4096;
        while (cap < bytes) cap *= 2;
        QS_HIP(hipHostMalloc(&R.buf[k], cap, hipHostMallocDefault));
        R.cap[k] = cap;
    }
    std::memcpy(R.buf[k], src, bytes);
    QS_HIP(hipMemcpyAsync(dst, R.buf[k], bytes, hipMemcpyHostToDevice, stream));
    QS_HIP(hipEventRecord(R.ev[k], stream));
    R.used[k] = true;
    return QS_OK;
}

namespace {

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

int check_state(const void* state, int nq) {
    QS_REQUIRE(state != nullptr, QS_E_ARG, "state pointer is null");
    QS_REQUIRE(nq >= 1 && nq <= SV_MAX_QUBITS, QS_E_RANGE, "nq=%d outside [1, %d]", nq,
               SV_MAX_QUBITS);
    return QS_OK;
}

// 2x2 of a one-qubit opcode of the statevector kernel (opt_kernel.cpp:257-347), with
// the parameters already decoded from their float32 words.
bool words_matrix(int op, double p0, double p1, double p2, double* m, bool* diag) {
    const double h = 0.7071067811865476;
    std::memset(m, 0, sizeof(double) * 8);
    auto set = [&](double a, double b, double c, double d, double e, double f, double g,
                   double k) {
        m[0] = a; m[1] = b; m[2] = c; m[3] = d; m[4] = e; m[5] = f; m[6] = g; m[7] = k;
    };
    *diag = false;
    switch (op) {
        case 1: set(h, 0, h, 0, h, 0, -h, 0); break;
        case 2: set(0, 0, 1, 0, 1, 0, 0, 0); break;
        case 3: set(0, 0, 0, -1, 0, 1, 0, 0); break;
        case 4: set(1, 0, 0, 0, 0, 0, -1, 0); *diag = true; break;
        case 5: set(1, 0, 0, 0, 0, 0, 0, 1); *diag = true; break;
        case 6: set(1, 0, 0, 0, 0, 0, h, h); *diag = true; break;
        case 7: set(1, 0, 0, 0, 0, 0, 0, -1); *diag = true; break;
        case 8: set(1, 0, 0, 0, 0, 0, h, -h); *diag = true; break;
        case 9: set(0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 0.5); break;
        case 10: set(0.5, -0.5, 0.5, 0.5, 0.5, 0.5, 0.5, -0.5); break;
        case 11: {
            const double c = std::cos(p0 * 0.5), s = std::sin(p0 * 0.5);
            set(c, 0, 0, -s, 0, -s, c, 0);
            break;
        }
        case 12: {
            const double c = std::cos(p0 * 0.5), s = std::sin(p0 * 0.5);
            set(c, 0, -s, 0, s, 0, c, 0);
            break;
        }
        case 13: {
            const double c = std::cos(p0 * 0.5), s = std::sin(p0 * 0.5);
            set(c, -s, 0, 0, 0, 0, c, s);
            *diag = true;
            break;
        }
        case 14: set(1, 0, 0, 0, 0, 0, std::cos(p0), std::sin(p0)); *diag = true; break;
        case 15:
            set(h, 0, -h * std::cos(p1), -h * std::sin(p1), h * std::cos(p0), h * std::sin(p0),
                h * std::cos(p0 + p1), h * std::sin(p0 + p1));
            break;
        case 16: {
            const double ct = std::cos(p0 * 0.5), st = std::sin(p0 * 0.5);
            set(ct, 0, -st * std::cos(p2), -st * std::sin(p2), st * std::cos(p1),
                st * std::sin(p1), ct * std::cos(p1 + p2), ct * std::sin(p1 + p2));
            break;
        }
        default: return false;
    }
    return true;
}

double f32_word(int32_t w) {
    float f;
    std::memcpy(&f, &w, sizeof(f));
    return (double)f;
}

// The plan's tables as one host blob laid out as sv_blob_layout says (the block-factor
// table at the end is written on the device and is not part of the upload).
std::vector<char> plan_blob(const SvPlan& plan, int nq, SvBlobLayout* layout) {
    const SvBlobLayout L = sv_blob_layout(plan, nq);
    std::vector<char> host(L.factors, 0);
    auto put = [&](size_t off, const void* src, size_t bytes) {
        if (bytes) std::memcpy(host.data() + off, src, bytes);
    };
    put(L.passes, plan.passes.data(), plan.passes.size() * sizeof(DevPass));
    put(L.stages, plan.stages.data(), plan.stages.size() * sizeof(DevStage));
    put(L.gates, plan.gates.data(), plan.gates.size() * sizeof(DevGate));
    put(L.phases, plan.phases.data(), plan.phases.size() * sizeof(DevPhase));
    put(L.terms, plan.terms.data(), plan.terms.size() * sizeof(DevTerm));
    put(L.tfactors, plan.thread_factors.data(), plan.thread_factors.size() * sizeof(double));
    *layout = L;
    return host;
}

// Uploads the plan's tables into a stream-ordered allocation (through the pinned staging
// ring, so the host blob may go as soon as the call returns) and launches it.
int run_plan(double2* state, int nq, const SvPlan& plan, hipStream_t stream) {
    if (plan.passes.empty()) return QS_OK;
    SvBlobLayout L;
    const std::vector<char> host = plan_blob(plan, nq, &L);
    void* dev = nullptr;
    QS_HIP(hipMallocAsync(&dev, L.total, stream));
    int st = h2d_copy_staged(dev, host.data(), host.size(), stream);
    if (st == QS_OK) st = sv_launch_plan(state, nq, plan, dev, stream);
    const hipError_t fe = hipFreeAsync(dev, stream);
    if (st) return st;
    QS_REQUIRE(fe == hipSuccess, QS_E_HIP, "hipFreeAsync failed: %s", hipGetErrorString(fe));
    return QS_OK;
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" {

int qs_abi_version(void) { return QS_ABI_VERSION; }

const char* qs_last_error(void) { return g_last_error.c_str(); }

int qs_sv_max_qubits(void) { return SV_MAX_QUBITS; }

int qs_init(int device) {
    clear_error();
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    QS_REQUIRE(e == hipSuccess && count > 0, QS_E_NODEV, "no HIP device visible (%s)",
               hipGetErrorString(e));
    QS_REQUIRE(device >= 0 && device < count, QS_E_NODEV, "device %d not in [0, %d)", device,
               count);
    QS_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    QS_HIP(hipGetDeviceProperties(&prop, device));
    QS_REQUIRE(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, QS_E_NODEV,
               "device %d is %s; this library is built for gfx950 only", device,
               prop.gcnArchName);
    return mps_scratch_setup(device);
}

int qs_sv_set_zero(void* state, int nq, void* stream) {
    clear_error();
    int st = check_state(state, nq);
    if (st) return st;
    return sv_fill_zero(static_cast<double2*>(state), nq, as_stream(stream));
}

int qs_sv_apply(void* state, int nq, const qs_gate* gates, int ng, void* stream) {
    clear_error();
    int st = check_state(state, nq);
    if (st) return st;
    QS_REQUIRE(ng >= 0, QS_E_ARG, "negative gate count");
    if (ng == 0) return QS_OK;
    QS_REQUIRE(gates != nullptr, QS_E_ARG, "gate array is null");
    SvPlan plan;
    st = sv_build_plan(nq, gates, ng, &plan);
    if (st) return st;
    return run_plan(static_cast<double2*>(state), nq, plan, as_stream(stream));
}

int qs_sv_apply_words(void* state, int nq, const int32_t* words, int num_gates, void* stream) {
    clear_error();
    int st = check_state(state, nq);
    if (st) return st;
    QS_REQUIRE(num_gates >= 0, QS_E_ARG, "negative gate count");
    if (num_gates == 0) return QS_OK;
    QS_REQUIRE(words != nullptr, QS_E_ARG, "gate_sequence is null");
    std::vector<qs_gate> gates;
    gates.reserve(num_gates);
    for (int i = 0; i < num_gates; ++i) {
        const int32_t* w = words + 8 * (size_t)i;
        const int op = w[0];
        qs_gate g;
        std::memset(&g, 0, sizeof(g));
        if (op == 0) continue;
        if (op == 17 || op == 18) {
            QS_REQUIRE(w[1] >= 0 && w[1] < nq && w[2] >= 0 && w[2] < nq && w[1] != w[2],
                       QS_E_ARG, "gate %d: bad qubits (%d, %d)", i, w[1], w[2]);
            if (op == 17) {          // CX: X on q1 controlled by q0
                g.kind = QS_GATE_U1;
                g.q[0] = w[2];
                g.ctrl_mask = 1ull << w[1];
                g.m[2] = 1.0;
                g.m[4] = 1.0;
            } else {                 // CZ
                g.kind = QS_GATE_DIAG2;
                g.q[0] = w[1];
                g.q[1] = w[2];
                g.m[0] = g.m[2] = g.m[4] = 1.0;
                g.m[6] = -1.0;
            }
            gates.push_back(g);
            continue;
        }
        double m[8];
        bool diag = false;
        QS_REQUIRE(words_matrix(op, f32_word(w[4]), f32_word(w[5]), f32_word(w[6]), m, &diag),
                   QS_E_ARG, "gate %d: unknown opcode %d", i, op);
        QS_REQUIRE(w[1] >= 0 && w[1] < nq, QS_E_ARG, "gate %d: qubit %d out of range", i, w[1]);
        g.q[0] = w[1];
        if (diag) {
            g.kind = QS_GATE_DIAG1;
            g.m[0] = m[0]; g.m[1] = m[1]; g.m[2] = m[6]; g.m[3] = m[7];
        } else {
            g.kind = QS_GATE_U1;
            std::memcpy(g.m, m, sizeof(m));
        }
        gates.push_back(g);
    }
    if (gates.empty()) return QS_OK;
    SvPlan plan;
    st = sv_build_plan(nq, gates.data(), (int)gates.size(), &plan);
    if (st) return st;
    return run_plan(static_cast<double2*>(state), nq, plan, as_stream(stream));
}

}  // extern "C"

// A plan built once on the host and kept on the device (qs_sv_plan_create).
struct qs_sv_plan {
    int nq = 0;
    int device = 0;
    qs::SvPlan plan;
    void* dev = nullptr;       // uploaded tables + block-factor table
    size_t bytes = 0;
    std::vector<hipFunction_t> spec;   // per-pass kernels after qs_sv_plan_specialize
};

extern "C" {

int qs_sv_plan_create(int nq, const qs_gate* gates, int ng, qs_sv_plan** plan_out) {
    clear_error();
    QS_REQUIRE(plan_out != nullptr, QS_E_ARG, "plan_out is null");
    *plan_out = nullptr;
    QS_REQUIRE(nq >= 1 && nq <= SV_MAX_QUBITS, QS_E_RANGE, "nq=%d outside [1, %d]", nq,
               SV_MAX_QUBITS);
    QS_REQUIRE(ng >= 0 && (ng == 0 || gates), QS_E_ARG, "bad gate array");
    qs_sv_plan* p = new qs_sv_plan();
    p->nq = nq;
    int st = sv_build_plan(nq, gates, ng, &p->plan);
    if (st == QS_OK && !p->plan.passes.empty()) {
        SvBlobLayout L;
        const std::vector<char> host = plan_blob(p->plan, nq, &L);
        hipError_t e = hipGetDevice(&p->device);
        if (e == hipSuccess) e = hipMalloc(&p->dev, L.total);
        if (e == hipSuccess) e = hipMemcpy(p->dev, host.data(), host.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            set_error("plan upload failed: %s", hipGetErrorString(e));
            st = QS_E_HIP;
        }
        p->bytes = L.total;
        // The block-factor tables depend on the plan alone: written once here, read by
        // every launch.
        if (st == QS_OK) st = sv_plan_block_factors(nq, p->plan, p->dev, nullptr);
        if (st == QS_OK) {
            e = hipStreamSynchronize(nullptr);
            if (e != hipSuccess) {
                set_error("block-factor tables failed: %s", hipGetErrorString(e));
                st = QS_E_HIP;
            }
        }
    }
    if (st != QS_OK) {
        if (p->dev) (void)hipFree(p->dev);
        delete p;
        return st;
    }
    *plan_out = p;
    return QS_OK;
}

int qs_sv_plan_launch(const qs_sv_plan* plan, void* state, void* stream) {
    clear_error();
    QS_REQUIRE(plan != nullptr, QS_E_ARG, "plan is null");
    int st = check_state(state, plan->nq);
    if (st) return st;
    if (plan->plan.passes.empty()) return QS_OK;
    int dev = -1;
    QS_HIP(hipGetDevice(&dev));
    QS_REQUIRE(dev == plan->device, QS_E_ARG, "plan lives on device %d, current device is %d",
               plan->device, dev);
    return sv_launch_plan(static_cast<double2*>(state), plan->nq, plan->plan, plan->dev,
                          as_stream(stream), plan->spec.empty() ? nullptr : plan->spec.data(),
                          false);
}

int qs_sv_plan_specialize(qs_sv_plan* plan, int* kernels_compiled) {
    clear_error();
    QS_REQUIRE(plan != nullptr, QS_E_ARG, "plan is null");
    if (kernels_compiled) *kernels_compiled = 0;
    if (plan->plan.passes.empty()) return QS_OK;
    int dev = -1;
    QS_HIP(hipGetDevice(&dev));
    QS_REQUIRE(dev == plan->device, QS_E_ARG, "plan lives on device %d, current device is %d",
               plan->device, dev);
    std::vector<hipFunction_t> fns;
    const int st = sv_jit_specialize(plan->plan, dev, &fns, kernels_compiled);
    if (st) return st;
    plan->spec = std::move(fns);
    return QS_OK;
}

int qs_sv_jit_source(int nq, const qs_gate* gates, int ng, int pass, char* buf, size_t cap,
                     size_t* needed) {
    clear_error();
    QS_REQUIRE(needed != nullptr, QS_E_ARG, "needed is null");
    QS_REQUIRE(nq >= 1 && nq <= SV_MAX_QUBITS, QS_E_RANGE, "nq=%d outside [1, %d]", nq,
               SV_MAX_QUBITS);
    QS_REQUIRE(ng >= 0 && (ng == 0 || gates), QS_E_ARG, "bad gate array");
    SvPlan plan;
    int st = sv_build_plan(nq, gates, ng, &plan);
    if (st) return st;
    QS_REQUIRE(pass >= 0 && pass < (int)plan.passes.size(), QS_E_RANGE,
               "pass %d outside [0, %zu)", pass, plan.passes.size());
    const std::string src = sv_jit_source(plan, pass);
    *needed = src.size() + 1;
    if (buf && cap > 0) {
        const size_t n = std::min(cap - 1, src.size());
        std::memcpy(buf, src.data(), n);
        buf[n] = '\0';
    }
    return QS_OK;
}

int qs_sv_jit_check(int nq, const qs_gate* gates, int ng, int* nkernels, size_t* code_bytes) {
    clear_error();
    QS_REQUIRE(nq >= 1 && nq <= SV_MAX_QUBITS, QS_E_RANGE, "nq=%d outside [1, %d]", nq,
               SV_MAX_QUBITS);
    QS_REQUIRE(ng >= 0 && (ng == 0 || gates), QS_E_ARG, "bad gate array");
    SvPlan plan;
    int st = sv_build_plan(nq, gates, ng, &plan);
    if (st) return st;
    return sv_jit_check(plan, nkernels, code_bytes);
}

int qs_sv_plan_info(const qs_sv_plan* plan, int* nq, int* npasses, size_t* device_bytes) {
    clear_error();
    QS_REQUIRE(plan != nullptr, QS_E_ARG, "plan is null");
    if (nq) *nq = plan->nq;
    if (npasses) *npasses = (int)plan->plan.passes.size();
    if (device_bytes) *device_bytes = plan->bytes;
    return QS_OK;
}

int qs_sv_plan_destroy(qs_sv_plan* plan) {
    clear_error();
    if (plan == nullptr) return QS_OK;
    hipError_t e = hipSuccess;
    if (plan->dev) e = hipFree(plan->dev);
    delete plan;
    QS_REQUIRE(e == hipSuccess, QS_E_HIP, "hipFree failed: %s", hipGetErrorString(e));
    return QS_OK;
}

int qs_sv_plan_passes(int nq, const qs_gate* gates, int ng, int* npasses) {
    clear_error();
    QS_REQUIRE(npasses != nullptr, QS_E_ARG, "npasses is null");
    QS_REQUIRE(nq >= 1 && nq <= SV_MAX_QUBITS, QS_E_RANGE, "nq=%d outside [1, %d]", nq,
               SV_MAX_QUBITS);
    QS_REQUIRE(ng >= 0 && (ng == 0 || gates), QS_E_ARG, "bad gate array");
    SvPlan plan;
    int st = sv_build_plan(nq, gates, ng, &plan);
    if (st) return st;
    *npasses = (int)plan.passes.size();
    return QS_OK;
}

int qs_sv_plan_json(int nq, const qs_gate* gates, int ng, char* buf, size_t cap,
                    size_t* needed) {
    clear_error();
    QS_REQUIRE(nq >= 1 && nq <= SV_MAX_QUBITS, QS_E_RANGE, "nq=%d outside [1, %d]", nq,
               SV_MAX_QUBITS);
    QS_REQUIRE(ng >= 0 && (ng == 0 || gates), QS_E_ARG, "bad gate array");
    SvPlan plan;
    int st = sv_build_plan(nq, gates, ng, &plan);
    if (st) return st;
    std::ostringstream os;
    os.precision(17);
    os << "{\"passes\":[";
    for (size_t p = 0; p < plan.passes.size(); ++p) {
        const DevPass& P = plan.passes[p];
        os << (p ? "," : "") << "{\"K\":" << P.K << ",\"R\":" << P.R << ",\"load_regs\":["
           << P.load.regs[0] << "," << P.load.regs[1] << "," << P.load.regs[2] << ","
           << P.load.regs[3] << "],\"tq\":[";
        for (int i = 0; i < P.K; ++i) os << (i ? "," : "") << P.tq[i];
        // the tile-index map: [src, dst, mask] per run of consecutive block qubits
        os << "],\"runs\":[";
        for (int k = 0; k < SV_TILE_BITS + 1; ++k)
            os << (k ? "," : "") << "[" << P.run_src[k] << "," << P.run_dst[k] << ","
               << P.run_mask[k] << "]";
        os << "],\"stages\":[";
        for (int s = 0; s < P.nstage; ++s) {
            const DevStage& S = plan.stages[P.stage0 + s];
            os << (s ? "," : "") << "{\"regs\":[";
            for (int i = 0; i < P.R; ++i) os << (i ? "," : "") << S.regs[i];
            os << "],\"gates\":[";
            for (int gi = S.g0; gi < S.g1; ++gi) {
                const DevGate& G = plan.gates[P.gate0 + gi];
                os << (gi > S.g0 ? "," : "") << "{\"kind\":" << G.kind << ",\"r0\":" << G.r0
                   << ",\"r1\":" << G.r1 << ",\"qa\":" << G.qa << ",\"qb\":" << G.qb
                   << ",\"cmask\":" << G.cmask << ",\"aux\":" << G.aux << ",\"m\":[";
                // a SIGN op's m holds its SignDesc words (printed under "phase")
                for (int k = 0; k < kDevGateWords; ++k)
                    os << (k ? "," : "") << (G.kind == DG_PHASE_SIGN || G.kind == DG_SIGN_REG ? 0.0 : G.m[k]);
                os << "]";
                if (G.kind == DG_PHASE || G.kind == DG_PHASE_SIGN || G.kind == DG_SIGN_REG ||
                    (G.kind >= DG_ROT && G.aux != 0)) {
                    const DevPhase& Q = plan.phases[P.phase0 + G.qa];
                    os << ",\"phase\":{\"sign\":" << Q.sign << ",\"slot\":" << Q.slot
                       << ",\"m0\":" << Q.m0 << ",\"mb\":[";
                    for (int k = 0; k < SV_THREAD_BITS; ++k)
                        os << (k ? "," : "") << Q.mb[k];
                    os << "],\"pp\":[";
                    for (int k = 0; k < SV_THREAD_BITS; ++k)
                        os << (k ? "," : "") << Q.pp[k];
                    os << "]"
                       << ",\"c\":[" << Q.c[0] << "," << Q.c[1] << "],\"kr\":[";
                    for (int k = 0; k < 4; ++k)
                        os << (k ? "," : "") << Q.kr[k][0] << "," << Q.kr[k][1];
                    os << "],\"crr\":[";
                    for (int k = 0; k < 6; ++k)
                        os << (k ? "," : "") << Q.crr[k][0] << "," << Q.crr[k][1];
                    os << "],\"table\":[";
                    const int nthr = 1 << (P.K - P.R);
                    if (Q.sign) {
                        for (int t = 0; t < nthr; ++t)
                            os << (t ? "," : "") << plan.thread_masks[Q.tab + t];
                    } else {
                        const size_t n0 = (size_t)Q.tab * 2 * PHASE_THREAD_VALUES;
                        const size_t n1 = n0 + (size_t)nthr * 2 * PHASE_THREAD_VALUES;
                        for (size_t k = n0; k < n1; ++k)
                            os << (k > n0 ? "," : "") << plan.thread_factors[k];
                    }
                    os << "],\"terms\":[";
                    const int n = Q.nt + Q.nb;
                    for (int k = 0; k < n; ++k) {
                        const DevTerm& T = plan.terms[P.term0 + (k < Q.nt ? Q.t0 + k : Q.b0 + k - Q.nt)];
                        os << (k ? "," : "") << "[" << T.cls << "," << T.a << "," << T.b << ","
                           << T.re << "," << T.im << "]";
                    }
                    os << "]}";
                }
                os << "}";
            }
            os << "]}";
        }
        os << "]}";
    }
    os << "]}";
    const std::string s = os.str();
    if (needed) *needed = s.size() + 1;
    if (buf && cap > s.size()) std::memcpy(buf, s.c_str(), s.size() + 1);
    return QS_OK;
}

int qs_sv_time_passes(int enable) {
    clear_error();
    return sv_time_passes(enable);
}

int qs_sv_pass_times(double* ms_out, int cap, int* count) {
    clear_error();
    QS_REQUIRE(count != nullptr && cap >= 0 && (cap == 0 || ms_out != nullptr), QS_E_ARG,
               "bad output arguments");
    return sv_pass_times(ms_out, cap, count);
}

int qs_sv_norm_sq(const void* state, int nq, double* out_host, void* stream) {
    clear_error();
    int st = check_state(state, nq);
    if (st) return st;
    QS_REQUIRE(out_host != nullptr, QS_E_ARG, "output pointer is null");
    return sv_norm_sq(static_cast<const double2*>(state), nq, out_host, as_stream(stream));
}

int qs_sv_scale(void* state, int nq, double re, double im, void* stream) {
    clear_error();
    int st = check_state(state, nq);
    if (st) return st;
    return sv_scale(static_cast<double2*>(state), nq, re, im, as_stream(stream));
}

int qs_sv_reset_qubit(void* state, int nq, int q, void* stream) {
    clear_error();
    int st = check_state(state, nq);
    if (st) return st;
    QS_REQUIRE(q >= 0 && q < nq, QS_E_ARG, "qubit %d out of range [0, %d)", q, nq);
    double2* s = static_cast<double2*>(state);
    hipStream_t hs = as_stream(stream);
    st = sv_project(s, nq, q, hs);
    if (st) return st;
    double nsq = 0.0;
    st = sv_norm_sq(s, nq, &nsq, hs);
    if (st) return st;
    const double nb = std::sqrt(nsq);
    if (nb < 1e-12) return sv_fill_zero(s, nq, hs);
    return sv_scale(s, nq, 1.0 / nb, 0.0, hs);
}

int qs_sv_probabilities(const void* state, int nq, double* probs, void* stream) {
    clear_error();
    int st = check_state(state, nq);
    if (st) return st;
    QS_REQUIRE(probs != nullptr, QS_E_ARG, "probs pointer is null");
    return sv_probs(static_cast<const double2*>(state), nq, probs, as_stream(stream));
}

size_t qs_sv_sample_workspace_bytes(int nq) {
    if (nq < 1 || nq > 34) return 0;
    return sv_sample_workspace_bytes(nq);
}

int qs_sv_sample(const void* state, int nq, const double* uniforms, int shots,
                 int64_t* indices, void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    int st = check_state(state, nq);
    if (st) return st;
    QS_REQUIRE(shots >= 0, QS_E_ARG, "negative shot count %d", shots);
    if (shots == 0) return QS_OK;
    QS_REQUIRE(uniforms != nullptr && indices != nullptr && workspace != nullptr, QS_E_ARG,
               "null uniforms, indices or workspace");
    return sv_sample(static_cast<const double2*>(state), nq, uniforms, shots, indices,
                     static_cast<double*>(workspace), workspace_bytes, as_stream(stream));
}

}  // extern "C"
