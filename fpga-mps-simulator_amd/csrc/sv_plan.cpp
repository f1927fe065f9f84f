// Host planner for the fused statevector passes.
//
// Input: an ordered gate list (qs_gate).  Output: passes; each pass names a tile of
// at most SV_TILE_BITS qubits (qubits 0..SV_LOW_BITS-1 always included) and a list of
// stages, each stage naming the SV_REG_BITS tile qubits held in registers and the
// gates applied under that register layout.
//
// Ordering rules: a gate may move ahead of a deferred gate only if they commute by
// construction -- disjoint qubits, or both diagonal.  Targets of non-diagonal gates
// must lie in the tile; diagonal gates and controls may use any qubit.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>

#include "qsim_internal.h"

namespace qs {

namespace {

inline int popcount64(uint64_t x) { return __builtin_popcountll(x); }

struct Item {
    HostGate g;
    bool diag;
    uint64_t touch;     // every qubit the gate reads: targets, diag qubits, controls
    uint64_t targets;   // qubits that must be tile qubits (non-diagonal targets)
};

void mat2_mul(const double* b, const double* a, double* out) {  // out = b @ a (2x2)
    double r[8];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            double re = 0, im = 0;
            for (int k = 0; k < 2; ++k) {
                const double br = b[2 * (2 * i + k)], bi = b[2 * (2 * i + k) + 1];
                const double ar = a[2 * (2 * k + j)], ai = a[2 * (2 * k + j) + 1];
                re += br * ar - bi * ai;
                im += br * ai + bi * ar;
            }
            r[2 * (2 * i + j)] = re;
            r[2 * (2 * i + j) + 1] = im;
        }
    std::memcpy(out, r, sizeof(r));
}

int check_gate(int nq, const qs_gate& g, int idx) {
    const uint64_t all = (nq >= 64) ? ~0ull : ((1ull << nq) - 1);
    QS_REQUIRE(g.kind >= QS_GATE_U1 && g.kind <= QS_GATE_DIAG2, QS_E_ARG,
               "gate %d: unknown kind %d", idx, g.kind);
    const bool two = (g.kind == QS_GATE_U2 || g.kind == QS_GATE_DIAG2);
    QS_REQUIRE(g.q[0] >= 0 && g.q[0] < nq, QS_E_ARG, "gate %d: qubit %d out of range [0,%d)",
               idx, g.q[0], nq);
    if (two) {
        QS_REQUIRE(g.q[1] >= 0 && g.q[1] < nq, QS_E_ARG,
                   "gate %d: qubit %d out of range [0,%d)", idx, g.q[1], nq);
        QS_REQUIRE(g.q[1] != g.q[0], QS_E_ARG, "gate %d: repeated qubit %d", idx, g.q[0]);
    }
    QS_REQUIRE((g.ctrl_mask & ~all) == 0, QS_E_ARG, "gate %d: control mask beyond %d qubits",
               idx, nq);
    uint64_t tmask = 1ull << g.q[0];
    if (two) tmask |= 1ull << g.q[1];
    QS_REQUIRE((g.ctrl_mask & tmask) == 0, QS_E_ARG, "gate %d: control overlaps a target", idx);
    for (int k = 0; k < 32; ++k)
        QS_REQUIRE(std::isfinite(g.m[k]), QS_E_ARG, "gate %d: non-finite matrix entry", idx);
    return QS_OK;
}

// Local position of qubit q in a tile given by its sorted qubit list.
int local_pos(const int* tq, int K, int q) {
    for (int i = 0; i < K; ++i)
        if (tq[i] == q) return i;
    return -1;
}

// ------------------------------------------------------------------ phase groups
typedef std::complex<double> cd;

struct PTerm {
    int qa, qb;            // global qubits; qb < 0: one-body; qa < 0: constant
    cd f;
};

bool unit_modulus(const double* m, int n) {
    for (int k = 0; k < n; ++k) {
        const double a = std::hypot(m[2 * k], m[2 * k + 1]);
        if (std::fabs(a - 1.0) > 1e-13) return false;
    }
    return true;
}

// Decomposes a unitary diagonal gate into phase factors; false if it has to be
// applied element by element (non-unit entries or more than one control).
bool phase_terms(const HostGate& g, std::vector<PTerm>& out) {
    auto c_of = [](const double* m, int k) { return cd(m[2 * k], m[2 * k + 1]); };
    int a, b;
    cd d00, d01, d10, d11;
    if (g.kind == QS_GATE_DIAG1) {
        if (!unit_modulus(g.m, 2)) return false;
        const int nc = __builtin_popcountll(g.cmask);
        if (nc == 0) {
            const cd d0 = c_of(g.m, 0), d1 = c_of(g.m, 1);
            out.push_back({-1, -1, d0});
            out.push_back({g.q0, -1, d1 * std::conj(d0)});
            return true;
        }
        if (nc != 1) return false;
        a = __builtin_ctzll(g.cmask);     // control is the MSB of a 2-qubit diagonal
        b = g.q0;
        d00 = 1.0; d01 = 1.0; d10 = c_of(g.m, 0); d11 = c_of(g.m, 1);
    } else {
        if (g.cmask != 0 || !unit_modulus(g.m, 4)) return false;
        a = g.q0;
        b = g.q1;
        d00 = c_of(g.m, 0); d01 = c_of(g.m, 1); d10 = c_of(g.m, 2); d11 = c_of(g.m, 3);
    }
    out.push_back({-1, -1, d00});
    out.push_back({a, -1, d10 * std::conj(d00)});
    out.push_back({b, -1, d01 * std::conj(d00)});
    out.push_back({a, b, d11 * d00 * std::conj(d10) * std::conj(d01)});
    return true;
}

bool is_one(const cd& z) { return z.real() == 1.0 && z.imag() == 0.0; }
bool is_sign(const cd& z) { return (z.real() == 1.0 || z.real() == -1.0) && z.imag() == 0.0; }

int pair_index(int r, int s) {          // r < s < 4
    static const int idx[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
    return idx[r][s];
}

// Builds one DevPhase record (plus its terms) for diagonal gates `diag` executed under
// register layout `regs` of pass P.
void emit_phase_group(const std::vector<const HostGate*>& diag, const int* regs, DevPass& P,
                      SvPlan* plan) {
    const int K = P.K, R = P.R;
    int ridx[SV_TILE_BITS], tidx[SV_TILE_BITS];
    int nt = 0;
    for (int p = 0; p < K; ++p) {
        ridx[p] = -1;
        tidx[p] = -1;
    }
    for (int i = 0; i < R; ++i) ridx[regs[i]] = i;
    for (int p = 0; p < K; ++p)
        if (ridx[p] < 0) tidx[p] = nt++;
    auto local = [&](int q) {
        for (int i = 0; i < K; ++i)
            if (P.tq[i] == q) return i;
        return -1;
    };
    cd c = 1.0, kr[4] = {1.0, 1.0, 1.0, 1.0}, crr[6] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
    std::vector<DevTerm> bt, tt;       // block-class and thread-class terms
    auto add = [&](std::vector<DevTerm>& v, int cls, int a, int b, const cd& f) {
        DevTerm t{};
        t.cls = cls;
        t.a = a;
        t.b = b;
        t.re = f.real();
        t.im = f.imag();
        v.push_back(t);
    };
    std::vector<PTerm> pts;
    for (const HostGate* g : diag) phase_terms(*g, pts);
    for (const PTerm& t : pts) {
        if (is_one(t.f)) continue;
        if (t.qa < 0) {
            c *= t.f;
            continue;
        }
        const int la = local(t.qa);
        if (t.qb < 0) {
            if (la < 0) add(bt, PT_BB, t.qa, -1, t.f);
            else if (ridx[la] >= 0) kr[ridx[la]] *= t.f;
            else add(tt, PT_T1, tidx[la], -1, t.f);
            continue;
        }
        int lb = local(t.qb);
        int qa = t.qa, qb = t.qb;
        int la2 = la;
        // order the pair so that register > thread > block in the first slot
        auto rank = [&](int l) { return l < 0 ? 0 : (ridx[l] >= 0 ? 2 : 1); };
        if (rank(lb) > rank(la2)) {
            std::swap(la2, lb);
            std::swap(qa, qb);
        }
        const int ra = rank(la2), rb = rank(lb);
        if (ra == 0) add(bt, PT_BB, qa, qb, t.f);                         // block x block
        else if (ra == 1 && rb == 0) add(bt, PT_BT, tidx[la2], qb, t.f);  // thread x block
        else if (ra == 1) add(tt, PT_TT, tidx[la2], tidx[lb], t.f);       // thread x thread
        else if (rb == 0) add(bt, PT_BR, ridx[la2], qb, t.f);             // register x block
        else if (rb == 1) add(tt, PT_TR, ridx[la2], tidx[lb], t.f);       // register x thread
        else {                                                             // register x register
            const int r = std::min(ridx[la2], ridx[lb]), s2 = std::max(ridx[la2], ridx[lb]);
            crr[pair_index(r, s2)] *= t.f;
        }
    }
    DevPhase ph{};
    bool sign = is_sign(c);
    for (int i = 0; i < 4; ++i) sign = sign && is_sign(kr[i]);
    for (int i = 0; i < 6; ++i) sign = sign && is_sign(crr[i]);
    for (const DevTerm& t : bt) sign = sign && is_sign(cd(t.re, t.im));
    for (const DevTerm& t : tt) sign = sign && is_sign(cd(t.re, t.im));
    ph.sign = sign ? 1 : 0;
    ph.c[0] = c.real();
    ph.c[1] = c.imag();
    for (int i = 0; i < 4; ++i) {
        ph.kr[i][0] = kr[i].real();
        ph.kr[i][1] = kr[i].imag();
    }
    for (int i = 0; i < 6; ++i) {
        ph.crr[i][0] = crr[i].real();
        ph.crr[i][1] = crr[i].imag();
    }
    ph.t0 = (int)plan->terms.size() - P.term0;
    ph.nt = (int)tt.size();
    plan->terms.insert(plan->terms.end(), tt.begin(), tt.end());
    ph.b0 = (int)plan->terms.size() - P.term0;
    ph.nb = (int)bt.size();
    plan->terms.insert(plan->terms.end(), bt.begin(), bt.end());
    ph.slot = bt.empty() ? -1 : P.nslot++;
    DevGate d{};
    d.kind = DG_PHASE;
    d.qa = (int)plan->phases.size() - P.phase0;
    d.qb = -1;
    plan->phases.push_back(ph);
    plan->gates.push_back(d);
}

void build_stages(const std::vector<const Item*>& inc, DevPass& P, SvPlan* plan) {
    const int K = P.K, R = P.R;
    int load[4] = {0, 0, 0, 0};
    for (int i = 0; i < R; ++i) load[i] = K - R + i;
    int cur[4] = {load[0], load[1], load[2], load[3]};
    P.stage0 = (int)plan->stages.size();
    P.gate0 = (int)plan->gates.size();
    P.phase0 = (int)plan->phases.size();
    P.term0 = (int)plan->terms.size();
    P.nslot = 0;
    int nstage = 0;
    bool open = false;
    DevStage st{};
    std::vector<const HostGate*> diag;   // pending unitary diagonal gates of the stage

    auto regmask_of = [&](const int* r) {
        int m = 0;
        for (int i = 0; i < R; ++i) m |= 1 << r[i];
        return m;
    };
    auto flush_diag = [&]() {
        if (!diag.empty()) emit_phase_group(diag, cur, P, plan);
        diag.clear();
    };
    auto open_stage = [&](const int* regs) {
        flush_diag();
        if (open) {
            st.g1 = (int)plan->gates.size() - P.gate0;
            plan->stages.push_back(st);
            ++nstage;
        }
        std::memset(&st, 0, sizeof(st));
        for (int i = 0; i < R; ++i) st.regs[i] = regs[i];
        st.g0 = (int)plan->gates.size() - P.gate0;
        open = true;
    };
    auto reg_index = [&](int lp) {
        for (int i = 0; i < R; ++i)
            if (cur[i] == lp) return i;
        return -1;
    };

    for (size_t n = 0; n < inc.size(); ++n) {
        const Item& it = *inc[n];
        const HostGate& g = it.g;
        if (it.diag) {
            if (!open) open_stage(cur);
            std::vector<PTerm> probe;
            if (phase_terms(g, probe)) {
                diag.push_back(&it.g);
                continue;
            }
            // General diagonal (non-unit entries or several controls): per amplitude.
            flush_diag();
            DevGate d{};
            d.cmask = g.cmask;
            d.qa = g.q0;
            d.qb = (g.kind == QS_GATE_DIAG2) ? g.q1 : -1;
            d.kind = (g.kind == QS_GATE_DIAG1) ? DG_DIAG1 : DG_DIAG2;
            std::memcpy(d.m, g.m, sizeof(double) * (g.kind == QS_GATE_DIAG1 ? 4 : 8));
            plan->gates.push_back(d);
            continue;
        }
        int need[2];
        int nneed = 0;
        need[nneed++] = local_pos(P.tq, K, g.q0);
        if (g.kind == QS_GATE_U2) need[nneed++] = local_pos(P.tq, K, g.q1);
        const int curmask = regmask_of(cur);
        int needmask = 0;
        for (int i = 0; i < nneed; ++i) needmask |= 1 << need[i];
        if (!open || (needmask & ~curmask)) {
            int nm = needmask;
            if (needmask & ~curmask) {
                // New register layout: this gate's targets, then targets of the next
                // non-diagonal gates that still fit, then previous registers.
                for (size_t k = n + 1; k < inc.size() && popcount64(nm) < R; ++k) {
                    const Item& nx = *inc[k];
                    if (nx.diag) continue;
                    int m2 = 1 << local_pos(P.tq, K, nx.g.q0);
                    if (nx.g.kind == QS_GATE_U2) m2 |= 1 << local_pos(P.tq, K, nx.g.q1);
                    if (popcount64((uint64_t)(nm | m2)) <= R) nm |= m2;
                }
                for (int i = 0; i < R && popcount64(nm) < R; ++i) nm |= 1 << cur[i];
                for (int p = K - 1; p >= 0 && popcount64(nm) < R; --p) nm |= 1 << p;
                // the pending diagonal group runs under the old layout
                flush_diag();
                int k = 0;
                for (int p = 0; p < K; ++p)
                    if ((nm >> p) & 1) cur[k++] = p;
            }
            open_stage(cur);
        }
        flush_diag();
        DevGate d{};
        d.cmask = g.cmask;
        d.qa = d.qb = -1;
        if (g.kind == QS_GATE_U1) {
            d.kind = DG_U1;
            d.r0 = reg_index(need[0]);
            // Merge with a directly preceding U1 on the same register bit and controls.
            if ((int)plan->gates.size() > P.gate0 + st.g0) {
                DevGate& prev = plan->gates.back();
                if (prev.kind == DG_U1 && prev.r0 == d.r0 && prev.cmask == d.cmask) {
                    mat2_mul(g.m, prev.m, prev.m);
                    continue;
                }
            }
            std::memcpy(d.m, g.m, sizeof(double) * 8);
        } else {
            d.kind = DG_U2;
            d.r0 = reg_index(need[0]);
            d.r1 = reg_index(need[1]);
            std::memcpy(d.m, g.m, sizeof(double) * 32);
        }
        plan->gates.push_back(d);
    }
    flush_diag();
    if (open) {
        st.g1 = (int)plan->gates.size() - P.gate0;
        plan->stages.push_back(st);
        ++nstage;
    }
    P.nstage = nstage;
    P.nphase = (int)plan->phases.size() - P.phase0;
    P.slot0 = plan->nslot_total;
    plan->nslot_total += P.nslot;
}

}  // namespace

int sv_build_plan(int nq, const qs_gate* gates, int ng, SvPlan* plan) {
    plan->passes.clear();
    plan->stages.clear();
    plan->gates.clear();
    plan->phases.clear();
    plan->terms.clear();
    plan->nslot_total = 0;
    std::vector<Item> items;
    items.reserve(ng);
    for (int i = 0; i < ng; ++i) {
        int st = check_gate(nq, gates[i], i);
        if (st != QS_OK) return st;
        Item it{};
        it.g.kind = gates[i].kind;
        it.g.q0 = gates[i].q[0];
        it.g.q1 = gates[i].q[1];
        it.g.cmask = gates[i].ctrl_mask;
        std::memcpy(it.g.m, gates[i].m, sizeof(it.g.m));
        const bool two = (it.g.kind == QS_GATE_U2 || it.g.kind == QS_GATE_DIAG2);
        it.diag = (it.g.kind == QS_GATE_DIAG1 || it.g.kind == QS_GATE_DIAG2);
        uint64_t q = 1ull << it.g.q0;
        if (two) q |= 1ull << it.g.q1;
        it.touch = q | it.g.cmask;
        it.targets = it.diag ? 0 : q;
        items.push_back(it);
    }

    const int K = std::min(SV_TILE_BITS, nq);
    const int R = std::min(SV_REG_BITS, K);
    const int L = std::min(SV_LOW_BITS, K);
    const uint64_t low = (L >= 64) ? ~0ull : ((1ull << L) - 1);

    std::vector<char> done(items.size(), 0);
    size_t remaining = items.size();
    while (remaining > 0) {
        DevPass P{};
        P.K = K;
        P.R = R;
        P.tile_is_all = (K == nq) ? 1 : 0;
        uint64_t tile = low;
        if (P.tile_is_all) tile = (nq >= 64) ? ~0ull : ((1ull << nq) - 1);
        uint64_t blocked_any = 0;   // qubits of deferred non-diagonal gates
        uint64_t blocked_nd = 0;    // qubits of deferred diagonal gates (block non-diagonal)
        std::vector<const Item*> inc;
        std::vector<size_t> inc_idx;
        for (size_t i = 0; i < items.size(); ++i) {
            if (done[i]) continue;
            const Item& it = items[i];
            bool defer = it.diag ? (it.touch & blocked_any) != 0
                                 : (it.touch & (blocked_any | blocked_nd)) != 0;
            if (!defer && (it.targets & ~tile)) {
                if (popcount64(tile | it.targets) <= K)
                    tile |= it.targets;
                else
                    defer = true;
            }
            if (defer) {
                if (it.diag)
                    blocked_nd |= it.touch;
                else
                    blocked_any |= it.touch;
                continue;
            }
            inc.push_back(&it);
            inc_idx.push_back(i);
        }
        // Fill the tile up to K qubits so every pass covers 2^K amplitudes per block.
        for (int q = 0; q < nq && popcount64(tile) < K; ++q) tile |= 1ull << q;
        P.tile_mask = tile;
        int k = 0;
        for (int q = 0; q < nq; ++q)
            if ((tile >> q) & 1) P.tq[k++] = q;
        build_stages(inc, P, plan);
        plan->passes.push_back(P);
        for (size_t i : inc_idx) done[i] = 1;
        remaining -= inc_idx.size();
        if (inc_idx.empty()) {
            set_error("planner made no progress (internal error)");
            return QS_E_ARG;
        }
    }
    return QS_OK;
}

}  // namespace qs

namespace qs {

SvBlobLayout sv_blob_layout(const SvPlan& plan, int nq) {
    SvBlobLayout L{};
    size_t off = 0;
    L.passes = off;
    off += sv_blob_offset(plan.passes.size() * sizeof(DevPass));
    L.stages = off;
    off += sv_blob_offset(plan.stages.size() * sizeof(DevStage));
    L.gates = off;
    off += sv_blob_offset(std::max<size_t>(plan.gates.size(), 1) * sizeof(DevGate));
    L.phases = off;
    off += sv_blob_offset(std::max<size_t>(plan.phases.size(), 1) * sizeof(DevPhase));
    L.terms = off;
    off += sv_blob_offset(std::max<size_t>(plan.terms.size(), 1) * sizeof(DevTerm));
    L.factors = off;
    size_t nblocks = 1;
    if (!plan.passes.empty() && nq > plan.passes[0].K) nblocks = (size_t)1 << (nq - plan.passes[0].K);
    off += sv_blob_offset((size_t)std::max(plan.nslot_total, 1) * nblocks * PHASE_BLOCK_VALUES *
                          2 * sizeof(double));
    L.total = off;
    return L;
}

}  // namespace qs
