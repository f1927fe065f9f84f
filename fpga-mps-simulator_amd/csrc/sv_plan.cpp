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

bool is_sign_cz(const HostGate& g) {
    // diag(1, 1, 1, -1) exactly: a CZ, applied as a conditional negation.
    const double* m = g.m;
    return g.kind == QS_GATE_DIAG2 && m[0] == 1.0 && m[1] == 0.0 && m[2] == 1.0 && m[3] == 0.0 &&
           m[4] == 1.0 && m[5] == 0.0 && m[6] == -1.0 && m[7] == 0.0;
}

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

void build_stages(const std::vector<const Item*>& inc, DevPass& P, SvPlan* plan) {
    const int K = P.K, R = P.R;
    int load[4] = {0, 0, 0, 0};
    for (int i = 0; i < R; ++i) load[i] = K - R + i;
    int cur[4] = {load[0], load[1], load[2], load[3]};
    P.stage0 = (int)plan->stages.size();
    P.gate0 = (int)plan->gates.size();
    int nstage = 0;
    bool open = false;
    DevStage st{};

    auto regmask_of = [&](const int* r) {
        int m = 0;
        for (int i = 0; i < R; ++i) m |= 1 << r[i];
        return m;
    };
    auto open_stage = [&](const int* regs) {
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
            // Merge with a directly preceding one-qubit diagonal on the same qubit.
            if (g.kind == QS_GATE_DIAG1 && (int)plan->gates.size() > P.gate0 + st.g0) {
                DevGate& prev = plan->gates.back();
                if (prev.kind == DG_DIAG1 && prev.qa == g.q0 && prev.cmask == g.cmask) {
                    double a[8] = {prev.m[0], prev.m[1], 0, 0, 0, 0, prev.m[2], prev.m[3]};
                    double b[8] = {g.m[0], g.m[1], 0, 0, 0, 0, g.m[2], g.m[3]};
                    double r[8];
                    mat2_mul(b, a, r);
                    prev.m[0] = r[0]; prev.m[1] = r[1]; prev.m[2] = r[6]; prev.m[3] = r[7];
                    continue;
                }
            }
            DevGate d{};
            d.cmask = g.cmask;
            d.qa = g.q0;
            d.qb = (g.kind == QS_GATE_DIAG2) ? g.q1 : -1;
            if (is_sign_cz(g)) {
                d.kind = DG_SIGN2;
            } else {
                d.kind = (g.kind == QS_GATE_DIAG1) ? DG_DIAG1 : DG_DIAG2;
                std::memcpy(d.m, g.m, sizeof(double) * (g.kind == QS_GATE_DIAG1 ? 4 : 8));
            }
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
                int k = 0;
                for (int p = 0; p < K; ++p)
                    if ((nm >> p) & 1) cur[k++] = p;
            }
            open_stage(cur);
        }
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
    if (open) {
        st.g1 = (int)plan->gates.size() - P.gate0;
        plan->stages.push_back(st);
        ++nstage;
    }
    P.nstage = nstage;
}

}  // namespace

int sv_build_plan(int nq, const qs_gate* gates, int ng, SvPlan* plan) {
    plan->passes.clear();
    plan->stages.clear();
    plan->gates.clear();
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
