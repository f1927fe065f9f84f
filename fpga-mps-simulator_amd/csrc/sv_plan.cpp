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

typedef std::complex<double> cd;

struct Item {
    HostGate g;
    bool diag;
    uint64_t touch;     // every qubit the gate reads: targets, diag qubits, controls
    uint64_t targets;   // qubits that must be tile qubits (non-diagonal targets)
    // Uncontrolled unitary one-qubit gates run as scaled rotations (split_rotation):
    // rot 1 or 2 is the form, t its parameter, post the factor on the |1> output and
    // scale the constant left over (collected per pass into one DG_SCALE op).
    int rot;
    double t;
    cd post, scale;
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

// One-qubit unitary U (row-major complex m[0..7]) written as
//   U = scale * diag(1, post) * Rn * diag(1, pre),
//   form 1 (|U00| >= |U10|): Rn = [[1, -t], [t, 1]], t = |U10|/|U00|, scale = g |U00|
//   form 2:                   Rn = [[t, -1], [1, t]], t = |U00|/|U10|, scale = g |U10|
// with g = U00/|U00| (U10/|U10| when U00 = 0), post = U10/(g |U10|) and
// pre = -U01/(g |U10|).  U00, U01 and U10 are reproduced to rounding by construction;
// U11 is compared, so a matrix that is not unitary to 1e-14 is refused (false), as is a
// diagonal one (U10 = 0), which has no rotation.  Rn costs two FMAs per output pair
// component instead of the eight of a complex 2x2 product, and 0 <= t <= 1.
bool split_rotation(const double* m, int& form, double& t, cd& pre, cd& post, cd& scale) {
    const cd u00(m[0], m[1]), u01(m[2], m[3]), u10(m[4], m[5]), u11(m[6], m[7]);
    const double c = std::abs(u00), s = std::abs(u10);
    if (!(s > 0.0) || !std::isfinite(c) || !std::isfinite(s)) return false;
    const cd g = c > 0.0 ? u00 / c : u10 / s;
    post = u10 / (g * s);
    pre = -u01 / (g * s);
    const cd r11 = g * post * pre * c;
    if (std::abs(r11 - u11) > 1e-14 || std::fabs(c * c + s * s - 1.0) > 1e-14 ||
        std::fabs(std::abs(pre) - 1.0) > 1e-14 || std::fabs(std::abs(post) - 1.0) > 1e-14)
        return false;
    if (c >= s) {
        form = 1;
        t = s / c;
        scale = g * c;
    } else {
        form = 2;
        t = c / s;
        scale = g * s;
    }
    return true;
}

// True for the SWAP matrix exactly (row-major complex 4x4).
bool is_swap(const double* m) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            const double want = (r == c && (r == 0 || r == 3)) || (r == 1 && c == 2) ||
                                        (r == 2 && c == 1)
                                    ? 1.0
                                    : 0.0;
            if (m[2 * (4 * r + c)] != want || m[2 * (4 * r + c) + 1] != 0.0) return false;
        }
    return true;
}

// ------------------------------------------------------------------ phase groups

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
    // Per-thread table of the thread/register classes.
    const int nthr = 1 << (K - R);
    auto thread_factors = [&](int t, cd& A, cd* k) {
        A = c;
        for (int i = 0; i < 4; ++i) k[i] = kr[i];
        for (const DevTerm& x : tt) {
            const cd f(x.re, x.im);
            const bool ba = (t >> x.a) & 1, bb = x.b >= 0 && ((t >> x.b) & 1);
            if (x.cls == PT_T1 && ba) A *= f;
            else if (x.cls == PT_TT && ba && bb) A *= f;
            else if (x.cls == PT_TR && bb) k[x.a] *= f;
        }
    };
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
    if (sign) {
        ph.tab = (int)plan->thread_masks.size();
        static const uint32_t col[4] = {0xAAAAu, 0xCCCCu, 0xF0F0u, 0xFF00u};
        static const int pr[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
        auto mask_of = [&](const cd& A, const cd* k) {
            uint32_t M = A.real() < 0.0 ? 0xFFFFu : 0u;
            for (int r = 0; r < R; ++r)
                if (k[r].real() < 0.0) M ^= col[r];
            for (int i = 0; i < 6; ++i)
                if (pr[i][1] < R && crr[i].real() < 0.0) M ^= col[pr[i][0]] & col[pr[i][1]];
            return M;
        };
        for (int t = 0; t < nthr; ++t) {
            cd A, k[4];
            thread_factors(t, A, k);
            plan->thread_masks.push_back(mask_of(A, k));
        }
        // The same masks as a descriptor: thread 0's mask, one XOR mask per thread bit
        // (one-body terms negate every register index, register x thread terms one
        // column) and the thread-pair terms.
        ph.m0 = plan->thread_masks[ph.tab];
        for (const DevTerm& x : tt) {
            if (x.cls == PT_T1) {
                ph.mb[x.a] ^= 0xFFFFu;
            } else if (x.cls == PT_TR) {
                ph.mb[x.b] ^= col[x.a];
            } else if (x.cls == PT_TT) {
                const int a = std::min(x.a, x.b), b = std::max(x.a, x.b);
                ph.pp[a] ^= 1u << b;
            }
        }
    } else {
        ph.tab = (int)(plan->thread_factors.size() / (2 * PHASE_THREAD_VALUES));
        for (int t = 0; t < nthr; ++t) {
            cd A, k[4];
            thread_factors(t, A, k);
            const cd v[PHASE_THREAD_VALUES] = {A, k[0], k[1], k[2], k[3]};
            for (const cd& z : v) {
                plan->thread_factors.push_back(z.real());
                plan->thread_factors.push_back(z.imag());
            }
        }
    }
    ph.t0 = (int)plan->terms.size() - P.term0;
    ph.nt = (int)tt.size();
    plan->terms.insert(plan->terms.end(), tt.begin(), tt.end());
    ph.b0 = (int)plan->terms.size() - P.term0;
    ph.nb = (int)bt.size();
    plan->terms.insert(plan->terms.end(), bt.begin(), bt.end());
    ph.slot = bt.empty() ? -1 : P.nslot++;
    DevGate d{};
    d.kind = sign ? DG_PHASE_SIGN : DG_PHASE;
    if (sign) {
        SignDesc sd{};
        sd.m0 = ph.m0;
        std::memcpy(sd.mb, ph.mb, sizeof(sd.mb));
        std::memcpy(sd.pp, ph.pp, sizeof(sd.pp));
        sd.slot = ph.slot;
        static_assert(sizeof(SignDesc) <= sizeof(d.m), "SignDesc must fit in DevGate::m");
        std::memcpy(d.m, &sd, sizeof(sd));
        bool uniform = ph.slot < 0;
        for (int b = 0; b < SV_THREAD_BITS; ++b) uniform = uniform && !ph.mb[b] && !ph.pp[b];
        if (uniform) {
            d.kind = DG_SIGN_REG;
            d.r0 = (int32_t)(ph.m0 & ((1u << (1 << P.R)) - 1u));   // 2^R register indices
        }
    }
    d.qa = (int)plan->phases.size() - P.phase0;
    d.qb = -1;
    if (!sign) {
        d.qb = ph.tab;
        d.aux = ph.slot;
        // which parts of the block-factor record are not all ones: bit 0 the thread-bit
        // factors (thread x block terms), bit 1 the register factors (register x block)
        for (const DevTerm& x : bt) {
            if (x.cls == PT_BT) d.r0 |= 1;
            if (x.cls == PT_BR) d.r0 |= 2;
        }
        for (int i = 0; i < 6; ++i) {
            d.m[2 * i] = ph.crr[i][0];
            d.m[2 * i + 1] = ph.crr[i][1];
        }
    }
    plan->phases.push_back(ph);
    plan->gates.push_back(d);
}

// ------------------------------------------------------------------ stage scheduler
//
// Within a pass every gate's targets are tile qubits, so the only constraint left is
// dependence order: gate j follows an earlier gate i when they share a qubit and are
// not both diagonal.  A stage fixes which R tile positions are register bits; the
// scheduler picks, stage by stage, the register set under which the most gates can
// run (following chains: a gate becomes ready once its predecessors ran), applies them
// in dependence order, and folds ready diagonal gates into one phase group just before
// the first gate that needs them.  Uncontrolled one-qubit gates that run back to back
// are merged into a DG_U1X record (one 2x2 matrix per register bit).
struct StageScheduler {
    const std::vector<const Item*>& inc;
    const DevPass& P;
    int K, R, L;
    int n;
    std::vector<int> need_mask;               // local-position mask of the targets
    std::vector<std::vector<int>> preds, succs;

    StageScheduler(const std::vector<const Item*>& items, const DevPass& pass)
        : inc(items), P(pass), K(pass.K), R(pass.R), L(std::min(SV_LOW_BITS, pass.K)) {
        n = (int)inc.size();
        need_mask.assign(n, 0);
        preds.assign(n, {});
        succs.assign(n, {});
        // Sparse dependences: per qubit, the last non-diagonal gate and the diagonal
        // gates seen since it.
        std::vector<int> last_nd(64, -1);
        std::vector<std::vector<int>> diag_since(64);
        for (int j = 0; j < n; ++j) {
            const Item& it = *inc[j];
            if (!it.diag) {
                need_mask[j] = 1 << local_pos(P.tq, K, it.g.q0);
                if (it.g.kind == QS_GATE_U2) need_mask[j] |= 1 << local_pos(P.tq, K, it.g.q1);
            }
            std::vector<int> pr;
            for (int q = 0; q < 64; ++q) {
                if (!((it.touch >> q) & 1ull)) continue;
                if (last_nd[q] >= 0) pr.push_back(last_nd[q]);
                if (!it.diag)
                    for (int d : diag_since[q]) pr.push_back(d);
            }
            std::sort(pr.begin(), pr.end());
            pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
            preds[j] = pr;
            for (int p : pr) succs[p].push_back(j);
            for (int q = 0; q < 64; ++q) {
                if (!((it.touch >> q) & 1ull)) continue;
                if (it.diag) {
                    diag_since[q].push_back(j);
                } else {
                    last_nd[q] = j;
                    diag_since[q].clear();
                }
            }
        }
    }

    struct State {
        std::vector<char> done;
        std::vector<int> rem;                 // unfinished predecessors
        int left_nd = 0;                      // non-diagonal gates not yet run
    };

    State initial() const {
        State s;
        s.done.assign(n, 0);
        s.rem.assign(n, 0);
        for (int j = 0; j < n; ++j) {
            s.rem[j] = (int)preds[j].size();
            if (!inc[j]->diag) ++s.left_nd;
        }
        return s;
    }

    void finish(State& s, int j) const {
        s.done[j] = 1;
        for (int k : succs[j]) --s.rem[k];
        if (!inc[j]->diag) --s.left_nd;
    }

    // Non-diagonal gate j can run now (its diagonal predecessors may still be pending
    // as long as they are themselves ready).
    bool runnable(const State& s, int j, bool* need_flush) const {
        bool flush = false;
        for (int p : preds[j]) {
            if (s.done[p]) continue;
            if (inc[p]->diag && s.rem[p] == 0) {
                flush = true;
                continue;
            }
            return false;
        }
        if (need_flush) *need_flush = flush;
        return true;
    }

    // Marks every ready diagonal gate done; returns their indices.
    std::vector<int> flush_ready(State& s) const {
        std::vector<int> out;
        for (int j = 0; j < n; ++j)
            if (!s.done[j] && inc[j]->diag && s.rem[j] == 0) out.push_back(j);
        for (int j : out) finish(s, j);
        return out;
    }

    // Runs every gate whose targets lie in `mask` until none is left; returns the
    // weighted count (two-qubit gates count twice).  `emit` receives the gate order.
    // Gates that need no pending diagonal run first, so one-qubit gates on distinct
    // register bits gather into one op and every flush serves as many gates as possible.
    template <class Emit>
    int run_stage(State& s, int mask, Emit&& emit) const {
        int count = 0;
        for (;;) {
            bool progress = false, blocked = false;
            for (int j = 0; j < n; ++j) {
                if (s.done[j] || inc[j]->diag || (need_mask[j] & ~mask)) continue;
                bool flush = false;
                if (!runnable(s, j, &flush)) continue;
                if (flush) {
                    blocked = true;
                    continue;
                }
                finish(s, j);
                emit(std::vector<int>(), j);
                count += (inc[j]->g.kind == QS_GATE_U2) ? 2 : 1;
                progress = true;
            }
            if (progress) continue;
            if (!blocked) break;
            emit(flush_ready(s), -1);
        }
        return count;
    }

    int score(const State& s, int mask) const {
        State t = s;
        const int full = run_stage(t, mask, [](const std::vector<int>&, int) {});
        // ready gates half inside the set make a partial set worth extending
        int partial = 0;
        for (int j = 0; j < n; ++j) {
            if (t.done[j] || inc[j]->diag) continue;
            if ((need_mask[j] & mask) && (need_mask[j] & ~mask) && runnable(t, j, nullptr))
                ++partial;
        }
        return 4 * full + partial;
    }

    // Register set of the next stage.  prefer_low: break ties toward the always-in-tile
    // low positions (keeps the high ones for the last stage, whose layout can then be
    // stored without a transpose).
    int choose(const State& s, bool prefer_low) const {
        int mask = 0;
        for (int slot = 0; slot < R; ++slot) {
            int best = -1, best_score = -1, best_pref = -1;
            for (int p = 0; p < K; ++p) {
                if ((mask >> p) & 1) continue;
                const int sc = score(s, mask | (1 << p));
                const bool low = p < L;
                const int pref = (prefer_low == low ? 2 * K : 0) + (low ? K - p : p);
                if (sc > best_score || (sc == best_score && pref > best_pref)) {
                    best = p;
                    best_score = sc;
                    best_pref = pref;
                }
            }
            mask |= 1 << best;
        }
        // Guarantee progress: the set must hold the targets of some runnable gate.
        State t = s;
        if (run_stage(t, mask, [](const std::vector<int>&, int) {}) == 0) {
            for (int j = 0; j < n; ++j) {
                if (s.done[j] || inc[j]->diag || !runnable(s, j, nullptr)) continue;
                int m2 = need_mask[j];
                for (int p = K - 1; p >= 0 && popcount64((uint64_t)m2) < R; --p)
                    if ((mask >> p) & 1) m2 |= 1 << p;
                for (int p = K - 1; p >= 0 && popcount64((uint64_t)m2) < R; --p) m2 |= 1 << p;
                return m2;
            }
        }
        return mask;
    }

    // Stage register masks of a whole schedule under one tie-break policy.
    std::vector<int> schedule(bool low_after_first) const {
        State s = initial();
        std::vector<int> masks;
        while (s.left_nd > 0) {
            const int m = choose(s, low_after_first && !masks.empty());
            run_stage(s, m, [](const std::vector<int>&, int) {});
            masks.push_back(m);
        }
        return masks;
    }

    int cost(const std::vector<int>& masks) const {
        if (masks.empty()) return 0;
        const int lowmask = (1 << L) - 1;
        int c = (int)masks.size() - 1;
        if (masks.front() & lowmask) ++c;
        if (masks.back() & lowmask) ++c;
        return c;
    }
};

// Workgroups are dispatched round-robin over the eight XCDs, so the low three bits of the
// tile index pick a tile's XCD.  Those bits go to the three highest block qubits: each XCD
// then streams its own eighth of the state.  Over the tile sets of the 28-qubit layered
// circuit a load/store probe ran 5 % faster on average (up to 14 %) than with the
// ascending map, and that circuit's step 2.5 % faster.  The remaining block qubits follow
// in ascending order.  P.bq lists the
// block qubits ascending on entry; on exit bit i of the tile index lands on P.bq[i], and
// the run tables describe the same map.
constexpr int kXcdBits = 3;

void set_runs(DevPass& P, int& nrun) {
    nrun = 0;
    for (int i = 0; i < P.nbq;) {
        int j = i;
        while (j + 1 < P.nbq && P.bq[j + 1] == P.bq[j] + 1) ++j;
        if (nrun < SV_TILE_BITS + 1) {
            P.run_src[nrun] = i;
            P.run_dst[nrun] = P.bq[i];
            P.run_mask[nrun] = (uint32_t)((1ull << (j - i + 1)) - 1);
        }
        ++nrun;
        i = j + 1;
    }
    for (int k = nrun; k < SV_TILE_BITS + 1; ++k) {
        P.run_src[k] = 0;
        P.run_dst[k] = 0;
        P.run_mask[k] = 0;
    }
}

void set_block_order(DevPass& P) {
    int nrun = 0;
    if (P.nbq > kXcdBits) {
        int ordered[SV_MAX_QUBITS];
        for (int i = 0; i < kXcdBits; ++i) ordered[i] = P.bq[P.nbq - kXcdBits + i];
        for (int i = kXcdBits; i < P.nbq; ++i) ordered[i] = P.bq[i - kXcdBits];
        int asc[SV_MAX_QUBITS];
        std::copy(P.bq, P.bq + P.nbq, asc);
        std::copy(ordered, ordered + P.nbq, P.bq);
        set_runs(P, nrun);
        if (nrun <= SV_TILE_BITS + 1) return;
        std::copy(asc, asc + P.nbq, P.bq);    // too many runs: keep the ascending map
    }
    set_runs(P, nrun);
}

void build_stages(const std::vector<const Item*>& inc, DevPass& P, SvPlan* plan) {
    const int K = P.K, R = P.R;
    P.stage0 = (int)plan->stages.size();
    P.gate0 = (int)plan->gates.size();
    P.phase0 = (int)plan->phases.size();
    P.term0 = (int)plan->terms.size();
    P.nslot = 0;

    StageScheduler S(inc, P);
    std::vector<int> masks = S.schedule(false);
    {
        std::vector<int> alt = S.schedule(true);
        if (S.cost(alt) < S.cost(masks)) masks.swap(alt);
    }
    if (masks.empty()) {   // only diagonal gates: one stage in the load layout
        int m = 0;
        for (int i = 0; i < R; ++i) m |= 1 << (K - R + i);
        masks.push_back(m);
    }

    StageScheduler::State st = S.initial();
    int cur[4] = {0, 0, 0, 0};
    // Constant factors of the rotations emitted since the last DG_SCALE.  A scaled
    // rotation grows amplitudes by the inverse of its factor (at most 1e3 sqrt(2)), so
    // the factor is applied whenever it falls below kScaleFloor (amplitudes within
    // about 1e33 of their true size, far from overflow, however many rotations a pass
    // holds) and once at the end of the pass.
    constexpr double kScaleFloor = 1e-30;
    cd scale = 1.0;
    auto emit_scale = [&]() {
        if (scale == cd(1.0, 0.0)) return;
        DevGate d{};
        d.kind = DG_SCALE;
        d.r0 = scale.imag() == 0.0 ? 1 : 0;
        d.qa = d.qb = -1;
        d.m[0] = scale.real();
        d.m[1] = scale.imag();
        plan->gates.push_back(d);
        scale = 1.0;
    };
    // Uncontrolled one-qubit gates gathered per register bit until another op is
    // emitted; three or four bits become one DG_U1X, fewer become DG_U1S ops.
    int u1_mask = 0;
    double u1_m[4][8];
    auto close_u1 = [&]() {
        if (!u1_mask) return;
        if (__builtin_popcount(u1_mask) >= 3 && R == 4) {
            DevGate d{};
            d.kind = DG_U1X;
            d.r0 = u1_mask;
            d.qa = d.qb = -1;
            for (int b = 0; b < 4; ++b) {
                if ((u1_mask >> b) & 1) {
                    std::memcpy(d.m + 8 * b, u1_m[b], sizeof(double) * 8);
                } else {
                    d.m[8 * b] = d.m[8 * b + 6] = 1.0;   // identity
                }
            }
            plan->gates.push_back(d);
        } else {
            for (int b = 0; b < 4; ++b) {
                if (!((u1_mask >> b) & 1)) continue;
                DevGate d{};
                d.kind = DG_U1S;
                d.r0 = b;
                d.qa = d.qb = -1;
                std::memcpy(d.m, u1_m[b], sizeof(double) * 8);
                plan->gates.push_back(d);
            }
        }
        u1_mask = 0;
    };
    auto reg_index = [&](int lp) {
        for (int i = 0; i < R; ++i)
            if (cur[i] == lp) return i;
        return -1;
    };
    // Index of the last rotation op that later first-form rotations with an output
    // factor may join (a DG_ROT of that variant or a DG_ROTS), valid while it is the
    // last op emitted.
    int open_rots = -1;
    // First op of the current stage (after its DG_LAYOUT) in plan->gates.
    size_t stage_first = plan->gates.size();
    auto emit_diagonals = [&](const std::vector<int>& idx) {
        if (idx.empty()) return;
        close_u1();
        std::vector<const HostGate*> group;
        std::vector<const HostGate*> general;
        for (int j : idx) {
            std::vector<PTerm> probe;
            if (phase_terms(inc[j]->g, probe))
                group.push_back(&inc[j]->g);
            else
                general.push_back(&inc[j]->g);
        }
        if (!group.empty()) {
            emit_phase_group(group, cur, P, plan);
            // A register-only sign group right after a rotation op of the same stage
            // rides on that op (its mask in aux, its phase record in qa): one dispatch
            // fewer.  The op then takes no further rotations, which would have to
            // precede the signs.
            const size_t n = plan->gates.size();
            if (n >= 2 && n - 2 >= stage_first && plan->gates[n - 1].kind == DG_SIGN_REG &&
                plan->gates[n - 2].kind >= DG_ROT && plan->gates[n - 2].aux == 0 &&
                plan->gates[n - 1].r0 != 0) {
                DevGate& rot = plan->gates[n - 2];
                rot.aux = plan->gates[n - 1].r0;
                rot.qa = plan->gates[n - 1].qa;
                plan->gates.pop_back();
                open_rots = -1;
            }
        }
        for (const HostGate* g : general) {   // non-unit entries or several controls
            DevGate d{};
            d.cmask = g->cmask;
            d.qa = g->q0;
            d.qb = (g->kind == QS_GATE_DIAG2) ? g->q1 : -1;
            d.kind = (g->kind == QS_GATE_DIAG1) ? DG_DIAG1 : DG_DIAG2;
            std::memcpy(d.m, g->m, sizeof(double) * (g->kind == QS_GATE_DIAG1 ? 4 : 8));
            plan->gates.push_back(d);
        }
    };
    auto emit_gate = [&](int j) {
        const HostGate& g = inc[j]->g;
        const int r0 = reg_index(local_pos(P.tq, K, g.q0));
        if (inc[j]->rot) {
            close_u1();
            // Second form as first form: [[t, -1], [1, t]] = t [[1, -1/t], [1/t, 1]],
            // with t joining the constant factor.  The rounding error stays that of the
            // second form; only amplitudes grow, by at most 1/kSecondFormMin per
            // rotation before the next DG_SCALE.
            constexpr double kSecondFormMin = 1e-3;
            int form = inc[j]->rot;
            double t = inc[j]->t;
            cd sc = inc[j]->scale;
            if (form == 2 && t >= kSecondFormMin) {
                form = 1;
                sc *= t;
                t = 1.0 / t;
            }
            const cd k = inc[j]->post;
            const bool post = k != cd(1.0, 0.0);
            const bool joins = form == 1 && post;
            DevGate* o = (joins && open_rots >= 0 &&
                          open_rots == (int)plan->gates.size() - 1 &&
                          plan->gates[open_rots].aux == 0)
                             ? &plan->gates[open_rots]
                             : nullptr;
            const int omask = !o ? 0 : o->kind >= DG_ROTS ? o->r0 : 1 << o->r0;
            if (o && !((omask >> r0) & 1)) {
                if (o->kind < DG_ROTS) {          // the single rotation becomes bit r0 of a run
                    const int b = o->r0;
                    const double tb = o->m[0], kr = o->m[2], ki = o->m[3];
                    std::fill(o->m, o->m + 12, 0.0);
                    o->m[b] = tb;
                    o->m[4 + 2 * b] = kr;
                    o->m[5 + 2 * b] = ki;
                    o->r1 = 0;
                }
                o->r0 = omask | (1 << r0);
                o->kind = DG_ROTS + o->r0 - 1;
                o->m[r0] = t;
                o->m[4 + 2 * r0] = k.real();
                o->m[5 + 2 * r0] = k.imag();
            } else {
                DevGate d{};
                d.r0 = r0;
                d.r1 = (form - 1) | (post ? 2 : 0);
                d.kind = DG_ROT + 4 * d.r0 + d.r1;
                d.qa = d.qb = -1;
                d.m[0] = t;
                d.m[2] = k.real();
                d.m[3] = k.imag();
                plan->gates.push_back(d);
                open_rots = joins ? (int)plan->gates.size() - 1 : -1;
            }
            scale *= sc;
            if (std::abs(scale) < kScaleFloor) emit_scale();
            return;
        }
        if (g.kind == QS_GATE_U1 && g.cmask == 0) {
            if ((u1_mask >> r0) & 1)
                mat2_mul(g.m, u1_m[r0], u1_m[r0]);
            else
                std::memcpy(u1_m[r0], g.m, sizeof(double) * 8);
            u1_mask |= 1 << r0;
            return;
        }
        close_u1();
        DevGate d{};
        d.cmask = g.cmask;
        d.qa = d.qb = -1;
        d.r0 = r0;
        if (g.kind == QS_GATE_U1) {
            d.kind = DG_U1;
            std::memcpy(d.m, g.m, sizeof(double) * 8);
        } else {
            d.kind = DG_U2;
            d.r1 = reg_index(local_pos(P.tq, K, g.q1));
            std::memcpy(d.m, g.m, sizeof(double) * 32);
            if (R == 4 && g.cmask == 0 && is_swap(g.m)) {
                static const int pair_of[4][4] = {
                    {-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
                d.kind = DG_SWAP;
                d.r0 = pair_of[d.r0][d.r1];
                d.r1 = 0;
                std::fill(d.m, d.m + 32, 0.0);
            }
        }
        plan->gates.push_back(d);
    };

    // HBM layouts: a layout whose register bits avoid the low positions keeps qubits
    // 0..L-1 on the lane index, so the first stage is loaded directly when it allows;
    // otherwise loads and stores use the default layout (the top R positions).
    const int L = std::min(SV_LOW_BITS, K);
    const int lowmask = (1 << L) - 1;
    int defmask = 0;
    for (int i = 0; i < R; ++i) defmask |= 1 << (K - R + i);
    const bool direct = (masks.front() & lowmask) == 0;
    const int loadmask = direct ? masks.front() : defmask;
    {
        int r[4] = {0, 0, 0, 0}, k = 0;
        for (int p = 0; p < K; ++p)
            if ((loadmask >> p) & 1) r[k++] = p;
        sv_fill_layout(P, r, &P.load);
    }
    auto emit_layout = [&](int mask) {
        close_u1();
        DevGate d{};
        d.kind = DG_LAYOUT;
        int r[4] = {0, 0, 0, 0}, k = 0;
        for (int p = 0; p < K; ++p)
            if ((mask >> p) & 1) r[k++] = p;
        d.r0 = r[0];
        d.r1 = r[1];
        d.qa = r[2];
        d.qb = r[3];
        DevLayout lay;
        sv_fill_layout(P, r, &lay);
        std::memcpy(d.m, &lay, sizeof(lay));
        plan->gates.push_back(d);
    };
    int curmask = loadmask;
    for (size_t si = 0; si < masks.size(); ++si) {
        int k = 0;
        for (int p = 0; p < K; ++p)
            if ((masks[si] >> p) & 1) cur[k++] = p;
        DevStage stg{};
        for (int i = 0; i < R; ++i) stg.regs[i] = cur[i];
        stg.g0 = (int)plan->gates.size() - P.gate0;
        if (masks[si] != curmask) emit_layout(masks[si]);
        stage_first = plan->gates.size();
        curmask = masks[si];
        S.run_stage(st, masks[si], [&](const std::vector<int>& diags, int j) {
            if (j < 0)
                emit_diagonals(diags);
            else
                emit_gate(j);
        });
        close_u1();
        if (si + 1 == masks.size()) {
            emit_diagonals(S.flush_ready(st));
            emit_scale();
            if ((curmask & lowmask) && curmask != defmask) emit_layout(defmask);
        }
        stg.g1 = (int)plan->gates.size() - P.gate0;
        plan->stages.push_back(stg);
    }
    P.nstage = (int)masks.size();
    P.nop = (int)plan->gates.size() - P.gate0;
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
    plan->thread_factors.clear();
    plan->thread_masks.clear();
    plan->nslot_total = 0;
    std::vector<Item> items;
    items.reserve(2 * (size_t)ng);
    // Last non-diagonal item on each qubit: a rotation's input factor diag(1, pre)
    // commutes with the diagonal gates since then, so it joins that item's output
    // factor when the item is a rotation on the same qubit.
    std::vector<int> last_nd(SV_MAX_QUBITS, -1);
    for (int i = 0; i < ng; ++i) {
        int st = check_gate(nq, gates[i], i);
        if (st != QS_OK) return st;
        Item it{};
        it.post = it.scale = 1.0;
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
        cd pre = 1.0;
        if (it.g.kind == QS_GATE_U1 && it.g.cmask == 0 &&
            split_rotation(it.g.m, it.rot, it.t, pre, it.post, it.scale) && pre != cd(1.0, 0.0)) {
            const int j = last_nd[it.g.q0];
            if (j >= 0 && items[j].rot) {
                items[j].post *= pre;
            } else {
                Item d{};
                d.g.kind = QS_GATE_DIAG1;
                d.g.q0 = it.g.q0;
                d.g.q1 = 0;
                d.g.cmask = 0;
                d.g.m[0] = 1.0;
                d.g.m[2] = pre.real();
                d.g.m[3] = pre.imag();
                d.diag = true;
                d.touch = q;
                d.targets = 0;
                d.post = d.scale = 1.0;
                items.push_back(d);
            }
        }
        if (!it.diag)
            for (int b = 0; b < nq; ++b)
                if ((it.touch >> b) & 1ull) last_nd[b] = (int)items.size();
        items.push_back(it);
    }

    const int K = std::min(SV_TILE_BITS, nq);
    const int R = std::min(SV_REG_BITS, K);
    const int L = std::min(SV_LOW_BITS, K);
    const uint64_t low = (L >= 64) ? ~0ull : ((1ull << L) - 1);

    std::vector<char> done(items.size(), 0);
    size_t remaining = items.size();
    // Indices of the items not yet placed, in gate order.
    std::vector<size_t> left(items.size());
    for (size_t i = 0; i < items.size(); ++i) left[i] = i;
    const uint64_t all_q = (nq >= 64) ? ~0ull : ((1ull << nq) - 1);
    // Items examined by the scans so far.  Once kSearchBudget is spent (long circuits of
    // many small passes, such as random gates on many qubits), the remaining passes take
    // the first-come tile, which costs one scan.
    constexpr size_t kSearchBudget = size_t(64) << 20;
    size_t scanned = 0;
    // One pass over the pending items `lst` in gate order: an item joins the pass unless
    // it must wait for a deferred item it does not commute with, or (non-diagonal) one of
    // its targets is not a tile qubit; with `extend` such a target joins the tile while
    // it has room.  Returns the number of non-diagonal items that join (the pass's work;
    // diagonal items cost little wherever they run), scanning at most `limit` items;
    // `inc` (when not null) receives the joining items and `tile` the final tile.
    auto scan = [&](const std::vector<size_t>& lst, uint64_t& tile, bool extend, size_t limit,
                    std::vector<size_t>* inc) {
        uint64_t blocked_any = 0;   // qubits of deferred non-diagonal gates
        uint64_t blocked_nd = 0;    // qubits of deferred diagonal gates (block non-diagonal)
        int work = 0;
        const size_t n = std::min(limit, lst.size());
        size_t k = 0;
        for (; k < n; ++k) {
            const Item& it = items[lst[k]];
            bool defer = it.diag ? (it.touch & blocked_any) != 0
                                 : (it.touch & (blocked_any | blocked_nd)) != 0;
            if (!defer && (it.targets & ~tile)) {
                if (extend && popcount64(tile | it.targets) <= K)
                    tile |= it.targets;
                else
                    defer = true;
            }
            if (defer) {
                if (it.diag)
                    blocked_nd |= it.touch;
                else
                    blocked_any |= it.touch;
                // Nothing later joins once every qubit is blocked for every item (or, when
                // only the work is counted, for every non-diagonal item).
                if ((blocked_any & all_q) == all_q ||
                    (!inc && ((blocked_any | blocked_nd) & all_q) == all_q))
                    break;
                continue;
            }
            if (!it.diag) ++work;
            if (inc) inc->push_back(lst[k]);
        }
        scanned += k;
        return work;
    };
    // Candidate tiles of a pass over `lst`, with the work each admits: the tile the scan
    // fills first-come (extend), a tile grown one qubit at a time by the most work, and
    // every tile of one or two runs of consecutive qubits above the low ones.  Candidates
    // are scored over the first kScoreItems pending items.
    struct Cand {
        int work;
        uint64_t tile;
        bool extend;
    };
    constexpr size_t kScoreItems = 8192;
    const int F = K - L;   // tile qubits above the low ones
    auto run = [](int a, int w) { return ((w >= 64 ? ~0ull : ((1ull << w) - 1)) << a); };
    auto candidates = [&](const std::vector<size_t>& lst, std::vector<Cand>* out) {
        out->clear();
        uint64_t t0 = low;
        const int w0 = scan(lst, t0, true, kScoreItems, nullptr);
        out->push_back({w0, t0, true});
        auto consider = [&](uint64_t cand) {
            uint64_t t = cand;
            out->push_back({scan(lst, t, false, kScoreItems, nullptr), cand, false});
        };
        uint64_t grow = low;
        while (popcount64(grow) < K) {
            int gbest = -1, gq = -1;
            for (int q = 0; q < nq; ++q) {
                if ((grow >> q) & 1) continue;
                uint64_t t = grow | (1ull << q);
                const int w = scan(lst, t, false, kScoreItems, nullptr);
                if (w > gbest) gbest = w, gq = q;
            }
            grow |= 1ull << gq;
        }
        consider(grow);
        for (int a = L; a + F <= nq; ++a) consider(low | run(a, F));
        for (int w1 = 1; w1 < F; ++w1)
            for (int a = L; a + w1 < nq; ++a)
                for (int b = a + w1 + 1; b + (F - w1) <= nq; ++b)
                    consider(low | run(a, w1) | run(b, F - w1));
    };
    std::vector<Cand> cands, next;
    std::vector<size_t> trial, rest;
    while (remaining > 0) {
        DevPass P{};
        P.K = K;
        P.R = R;
        P.tile_is_all = (K == nq) ? 1 : 0;
        uint64_t tile = low;
        if (P.tile_is_all) tile = (nq >= 64) ? ~0ull : ((1ull << nq) - 1);
        bool extend = true;
        if (!P.tile_is_all && scanned < kSearchBudget) {
            // The tile whose pass admits the most non-diagonal gates together with the best
            // pass after it, over the kLookahead candidates of most work (ties: the earlier
            // candidate, the first-come tile first).  Diagonal gates run under any tile, so
            // in a layered circuit of one-qubit gates and nearest-neighbour diagonals a
            // window of consecutive qubits advances a light cone: the 28-qubit benchmark
            // takes 17 passes instead of the 31 of first-come tiles, whose windows chased
            // single chains.
            constexpr int kLookahead = 4;
            candidates(left, &cands);
            std::stable_sort(cands.begin(), cands.end(),
                             [](const Cand& x, const Cand& y) { return x.work > y.work; });
            int best = -1;
            for (int c = 0; c < (int)cands.size() && c < kLookahead; ++c) {
                uint64_t t = cands[c].tile;
                trial.clear();
                scan(left, t, cands[c].extend, left.size(), &trial);
                rest.clear();
                size_t k = 0;
                for (size_t i : left) {
                    if (k < trial.size() && trial[k] == i) ++k;
                    else rest.push_back(i);
                }
                int after = 0;
                if (!rest.empty()) {
                    candidates(rest, &next);
                    for (const Cand& x : next) after = std::max(after, x.work);
                }
                if (cands[c].work + after > best) {
                    best = cands[c].work + after;
                    tile = cands[c].tile;
                    extend = cands[c].extend;
                }
            }
        }
        std::vector<size_t> inc_idx;
        scan(left, tile, extend, left.size(), &inc_idx);
        std::vector<const Item*> inc;
        inc.reserve(inc_idx.size());
        for (size_t i : inc_idx) inc.push_back(&items[i]);
        // Fill the tile up to K qubits so every pass covers 2^K amplitudes per block.
        for (int q = 0; q < nq && popcount64(tile) < K; ++q) tile |= 1ull << q;
        P.tile_mask = tile;
        int k = 0;
        for (int q = 0; q < nq; ++q)
            if ((tile >> q) & 1) P.tq[k++] = q;
        P.nbq = 0;
        if (!P.tile_is_all)
            for (int q = 0; q < nq; ++q)
                if (!((tile >> q) & 1)) P.bq[P.nbq++] = q;
        set_block_order(P);
        build_stages(inc, P, plan);
        plan->passes.push_back(P);
        for (size_t i : inc_idx) done[i] = 1;
        remaining -= inc_idx.size();
        left.erase(std::remove_if(left.begin(), left.end(), [&](size_t i) { return done[i] != 0; }),
                   left.end());
        if (inc_idx.empty()) {
            set_error("planner made no progress (internal error)");
            return QS_E_ARG;
        }
    }
    // DG_ROTS records were built with k_b at (m[4 + 2b], m[5 + 2b]); the kernels apply
    // the output factors as one diagonal over the register index, f_j = prod of k_b
    // over the bits b of j & mask, stored at (m[2 + 2j], m[3 + 2j]) for j = 1..15.
    for (DevGate& G : plan->gates) {
        if (G.kind < DG_ROTS || G.kind >= DG_ROTS + 15) continue;
        const int mask = G.r0;
        cd k[4];
        for (int b = 0; b < 4; ++b) k[b] = cd(G.m[4 + 2 * b], G.m[5 + 2 * b]);
        for (int j = 1; j < 16; ++j) {
            cd f(1.0, 0.0);
            for (int b = 0; b < 4; ++b)
                if (((j & mask) >> b) & 1) f *= k[b];
            G.m[2 + 2 * j] = f.real();
            G.m[3 + 2 * j] = f.imag();
        }
    }
    return QS_OK;
}

}  // namespace qs

namespace qs {

void sv_fill_layout(const DevPass& P, const int* regs, DevLayout* out) {
    std::memset(out, 0, sizeof(*out));
    int regmask = 0;
    for (int i = 0; i < P.R; ++i) {
        out->regs[i] = regs[i];
        out->rq[i] = P.tq[regs[i]];
        regmask |= 1 << regs[i];
    }
    int b = 0;
    for (int p = 0; p < P.K; ++p) {
        if ((regmask >> p) & 1) continue;
        out->tpos[b] = p;
        out->tq[b] = P.tq[p];
        ++b;
    }
}

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
    L.tfactors = off;
    off += sv_blob_offset(std::max<size_t>(plan.thread_factors.size(), 1) * sizeof(double));
    L.factors = off;
    size_t nblocks = 1;
    if (!plan.passes.empty() && nq > plan.passes[0].K) nblocks = (size_t)1 << (nq - plan.passes[0].K);
    off += sv_blob_offset((size_t)std::max(plan.nslot_total, 1) * nblocks * PHASE_BLOCK_VALUES *
                          2 * sizeof(double));
    L.total = off;
    return L;
}

}  // namespace qs
