"""Replay a fused-pass plan (qs_sv_plan_json) with NumPy, as the pass kernel reads it.

Test helper: it resolves every register bit through the stage's register list
and the pass's tile qubit list, exactly like sv_pass_kernel, and applies the
gates to a dense state.  A planner bug (wrong order, wrong register index,
target outside the tile) shows up as a mismatch against the oracle.
"""

import numpy as np

(DG_U1, DG_U2, DG_DIAG1, DG_DIAG2, DG_U1X, DG_PHASE, DG_LAYOUT, DG_PHASE_SIGN,
 DG_U1S, DG_SWAP) = range(10)
DG_SCALE = 10
DG_SIGN_REG = 11     # sign group on register bits only: mask in r0
DG_ROT = 16          # DG_ROT + 4 * register bit + variant
DG_ROTS = 32         # DG_ROTS + mask - 1: first-form rotations with factors on mask's bits
SECOND_FORM_MIN = 1e-3
PT_BB, PT_BT, PT_BR, PT_T1, PT_TT, PT_TR = 0, 1, 2, 3, 4, 5
PAIRS = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def _cm(vals, n):
    v = np.asarray(vals[:2 * n], dtype=np.float64)
    return v[0::2] + 1j * v[1::2]


def _ctrl_ok(idx, cmask):
    return (idx & cmask) == cmask


def _apply_1q(st, idx, sel, t, m):
    lo = sel & (((idx >> t) & 1) == 0)
    i0 = idx[lo]
    i1 = i0 | (1 << t)
    a0, a1 = st[i0].copy(), st[i1].copy()
    st[i0] = m[0, 0] * a0 + m[0, 1] * a1
    st[i1] = m[1, 0] * a0 + m[1, 1] * a1


def replay(plan: dict, nq: int, state: np.ndarray) -> np.ndarray:
    st = state.copy()
    idx = np.arange(st.size, dtype=np.int64)
    for p in plan['passes']:
        tq = p['tq']
        R = p['R']
        assert len(tq) == p['K'] and sorted(tq) == tq
        tile = 0
        for q in tq:
            tile |= 1 << q
        low = min(4, p['K'])
        layout = list(p['load_regs'][:R])
        # HBM layouts keep the low positions on the lane index (or are the default)
        default = list(range(p['K'] - R, p['K']))
        assert all(r >= low for r in layout) or layout == default
        for si, s in enumerate(p['stages']):
            regs = s['regs']
            assert len(regs) == R and sorted(regs) == regs
            for gi, g in enumerate(s['gates']):
                if g['kind'] == DG_LAYOUT:
                    layout = [g['r0'], g['r1'], g['qa'], g['qb']][:R]
                    assert sorted(layout) == layout
                    continue
                # every other op runs under the stage's register layout
                assert layout == regs, (si, gi, layout, regs)
                cm = int(g['cmask'])
                sel = _ctrl_ok(idx, cm)
                if g['kind'] == DG_U1:
                    _apply_1q(st, idx, sel, tq[regs[g['r0']]], _cm(g['m'], 4).reshape(2, 2))
                elif g['kind'] == DG_U1X:
                    # the kernel applies all four matrices; bits outside r0 hold identity
                    assert cm == 0 and R == 4 and bin(g['r0']).count('1') >= 3
                    allm = _cm(g['m'], 16)
                    for b in range(len(regs)):
                        mb = allm[4 * b:4 * b + 4].reshape(2, 2)
                        if not (g['r0'] >> b) & 1:
                            assert np.array_equal(mb, np.eye(2))
                        _apply_1q(st, idx, sel, tq[regs[b]], mb)
                elif DG_ROT <= g['kind'] < DG_ROT + 16:
                    # scaled rotation: form A [[1, -t], [t, 1]], form B [[t, -1], [1, t]];
                    # r1 bit 1: second row times k; bit and variant are also in the kind
                    assert cm == 0 and 0 <= g['r0'] < R
                    assert g['kind'] == DG_ROT + 4 * g['r0'] + g['r1']
                    t = g['m'][0]
                    m = (np.array([[t, -1], [1, t]], dtype=complex) if g['r1'] & 1
                         else np.array([[1, -t], [t, 1]], dtype=complex))
                    # the planner writes second-form rotations in the first form unless
                    # t is tiny, so first-form t reaches 1 / SECOND_FORM_MIN
                    assert (0.0 <= t < SECOND_FORM_MIN if g['r1'] & 1
                            else 0.0 <= t <= 1.0 / SECOND_FORM_MIN)
                    if g['r1'] & 2:
                        m[1] *= g['m'][2] + 1j * g['m'][3]
                    _apply_1q(st, idx, sel, tq[regs[g['r0']]], m)
                elif DG_ROTS <= g['kind'] < DG_ROTS + 15:
                    mask = g['r0']
                    assert cm == 0 and g['kind'] == DG_ROTS + mask - 1 and 0 < mask < (1 << R)
                    assert bin(mask).count('1') >= 2, 'a lone rotation stays a DG_ROT'
                    # the rotations, then their output factors as one diagonal over the
                    # register index j: f_j = (m[2 + 2j], m[3 + 2j]), the product of the
                    # single-bit factors f_(2^b) over the bits of j & mask
                    f = {j: g['m'][2 + 2 * j] + 1j * g['m'][3 + 2 * j] for j in range(1, 16)}
                    for b in range(R):
                        if (mask >> b) & 1:
                            t = g['m'][b]
                            assert 0.0 <= t <= 1.0 / SECOND_FORM_MIN
                            m = np.array([[1, -t], [t, 1]], dtype=complex)
                            _apply_1q(st, idx, sel, tq[regs[b]], m)
                            assert (tile >> tq[regs[b]]) & 1
                    reg = np.zeros_like(idx)
                    for b in range(R):
                        reg |= ((idx >> tq[regs[b]]) & 1) << b
                    for j in range(1, 1 << R):
                        if j & mask:
                            want = np.prod([f[1 << b] for b in range(R) if (j & mask) >> b & 1])
                            assert abs(f[j] - want) <= 1e-15 * (1 + abs(want)), (j, f[j], want)
                            st[sel & (reg == j)] *= f[j]
                elif g['kind'] == DG_SCALE:
                    f = g['m'][0] + 1j * g['m'][1]
                    assert (g['r0'] == 1) == (g['m'][1] == 0.0)
                    st *= f
                elif g['kind'] == DG_U1S:
                    assert cm == 0 and 0 <= g['r0'] < R
                    _apply_1q(st, idx, sel, tq[regs[g['r0']]], _cm(g['m'], 4).reshape(2, 2))
                elif g['kind'] == DG_U2:
                    ta = tq[regs[g['r0']]]
                    tb = tq[regs[g['r1']]]
                    m = _cm(g['m'], 16).reshape(4, 4)
                    base = idx[sel & (((idx >> ta) & 1) == 0) & (((idx >> tb) & 1) == 0)]
                    ii = [base, base | (1 << tb), base | (1 << ta), base | (1 << ta) | (1 << tb)]
                    v = np.stack([st[i] for i in ii])
                    out = m @ v
                    for k in range(4):
                        st[ii[k]] = out[k]
                elif g['kind'] == DG_SWAP:
                    assert cm == 0 and R == 4
                    ba, bb = PAIRS[g['r0']]
                    qa, qb = tq[regs[ba]], tq[regs[bb]]
                    assert (tile >> qa) & 1 and (tile >> qb) & 1
                    i10 = idx[(((idx >> qa) & 1) == 1) & (((idx >> qb) & 1) == 0)]
                    i01 = i10 ^ (1 << qa) ^ (1 << qb)
                    st[i10], st[i01] = st[i01].copy(), st[i10].copy()
                elif g['kind'] == DG_SIGN_REG:
                    st *= register_signs(g['phase'], g['r0'], tq, regs, idx)
                elif g['kind'] in (DG_PHASE, DG_PHASE_SIGN):
                    assert (g['kind'] == DG_PHASE_SIGN) == bool(g['phase']['sign'])
                    if g['kind'] == DG_PHASE:
                        # the kernel skips the block factors whose flag is clear
                        classes = {c for c, *_ in g['phase']['terms']}
                        assert bool(g['r0'] & 1) == (PT_BT in classes), 'thread x block flag'
                        assert bool(g['r0'] & 2) == (PT_BR in classes), 'register x block flag'
                    st *= phase_factor(g['phase'], tq, regs, idx)
                elif g['kind'] == DG_DIAG1:
                    d = _cm(g['m'], 2)
                    b = (idx >> g['qa']) & 1
                    st[sel] *= d[b[sel]]
                elif g['kind'] == DG_DIAG2:
                    d = _cm(g['m'], 4)
                    k = 2 * ((idx >> g['qa']) & 1) + ((idx >> g['qb']) & 1)
                    st[sel] *= d[k[sel]]
                else:
                    raise AssertionError(f"unknown kind {g['kind']}")
                # every non-diagonal target must be a tile qubit
                # a register-only sign group riding on a rotation op, applied after it
                if g['kind'] >= DG_ROT and g.get('aux', 0):
                    st *= register_signs(g['phase'], g['aux'], tq, regs, idx)
                if g['kind'] in (DG_U1, DG_U2, DG_U1S) or DG_ROT <= g['kind'] < DG_ROTS:
                    assert (tile >> tq[regs[g['r0']]]) & 1
                if g['kind'] == DG_U1:
                    assert cm != 0, 'uncontrolled one-qubit gates travel as U1X, U1S or ROT'
        assert all(r >= low for r in layout) or layout == default, layout
    return st


def register_signs(ph, mask, tq, regs, idx):
    """Factor of a sign group on register bits only, as the kernel reads it: amplitude j
    of the register block is negated when bit j of `mask` is set.  Checked against the
    group's own terms."""
    assert ph['sign'] and ph['slot'] < 0 and not any(ph['mb']) and not any(ph['pp'])
    f = phase_factor(ph, tq, regs, idx)
    j = np.zeros_like(idx)
    for b in range(len(regs)):
        j |= ((idx >> tq[regs[b]]) & 1) << b
    fk = np.where((mask >> j) & 1, -1.0, 1.0)
    assert np.allclose(f, fk, rtol=0, atol=1e-12), 'register sign mask'
    return fk


def phase_factor(ph, tq, regs, idx):
    """Per-amplitude factor of a phase group, resolving each term class to qubits.

    Also rebuilds the factor the way the kernel does (per-thread table for the
    thread/register classes, block terms on top) and checks that both agree.
    """
    f = _phase_factor_terms(ph, tq, regs, idx)
    g = _phase_factor_table(ph, tq, regs, idx)
    assert np.allclose(f, g, rtol=0, atol=1e-12), np.max(np.abs(f - g))
    return f


def _thread_and_reg_index(tq, regs, idx):
    K = len(tq)
    thread_pos = [p for p in range(K) if p not in regs]
    t = np.zeros(idx.size, dtype=np.int64)
    for b, p in enumerate(thread_pos):
        t |= ((idx >> tq[p]) & 1) << b
    j = np.zeros(idx.size, dtype=np.int64)
    for r, p in enumerate(regs):
        j |= ((idx >> tq[p]) & 1) << r
    return t, j, thread_pos


def _sign_mask_from_descriptor(ph, t):
    """Thread t's negation mask as sv_pass_kernel's sign_mask() computes it."""
    M = int(ph['m0'])
    for b, mb in enumerate(ph['mb']):
        if (t >> b) & 1:
            M ^= int(mb)
    par = 0
    for a, row in enumerate(ph['pp']):
        par ^= ((t >> a) & 1) & bin(int(row) & t).count('1')
    return M ^ (0xFFFF if par & 1 else 0)


def _phase_factor_table(ph, tq, regs, idx):
    t, j, thread_pos = _thread_and_reg_index(tq, regs, idx)
    bit = lambda q: (idx >> q) & 1
    tab = np.asarray(ph['table'])
    R = len(regs)
    if ph['sign']:
        # the kernel rebuilds the table from the group's descriptor
        for tt in range(tab.size):
            assert _sign_mask_from_descriptor(ph, tt) == int(tab[tt]), (tt, ph['m0'])
        M = tab.astype(np.int64)[t]
        f = np.where((M >> j) & 1 == 1, -1.0 + 0j, 1.0 + 0j)
    else:
        v = tab.reshape(-1, 5, 2)
        e = v[..., 0] + 1j * v[..., 1]           # (threads, 5): A, kappa_0..3
        f = e[t, 0].copy()
        for r in range(R):
            f = np.where((j >> r) & 1 == 1, f * e[t, 1 + r], f)
        for k, (r, s2) in enumerate(PAIRS):
            if s2 < R:
                c = complex(ph['crr'][2 * k], ph['crr'][2 * k + 1])
                f = np.where(((j >> r) & (j >> s2) & 1) == 1, f * c, f)
    # block classes, as the prologue folds them
    for cls, a, b, re, im in ph['terms']:
        z = complex(re, im)
        if cls == PT_BB:
            on = bit(a) & (bit(b) if b >= 0 else 1)
        elif cls == PT_BT:
            on = bit(tq[thread_pos[a]]) & bit(b)
        elif cls == PT_BR:
            on = bit(tq[regs[a]]) & bit(b)
        else:
            continue
        f = np.where(on == 1, f * z, f)
    return f


def _phase_factor_terms(ph, tq, regs, idx):
    K = len(tq)
    thread_pos = [p for p in range(K) if p not in regs]      # thread bit i -> local position
    bit = lambda q: (idx >> q) & 1
    tq_thread = lambda a: tq[thread_pos[a]]
    tq_reg = lambda a: tq[regs[a]]
    f = np.full(idx.size, complex(*ph['c']))
    vals = []
    for r in range(len(regs)):
        vals.append((bit(tq_reg(r)), complex(ph['kr'][2 * r], ph['kr'][2 * r + 1])))
    for k, (r, s2) in enumerate(PAIRS):
        if s2 < len(regs):
            vals.append((bit(tq_reg(r)) & bit(tq_reg(s2)),
                         complex(ph['crr'][2 * k], ph['crr'][2 * k + 1])))
    for cls, a, b, re, im in ph['terms']:
        if cls == PT_BB:
            on = bit(a) & (bit(b) if b >= 0 else 1)
        elif cls == PT_BT:
            on = bit(tq_thread(a)) & bit(b)
        elif cls == PT_BR:
            on = bit(tq_reg(a)) & bit(b)
        elif cls == PT_T1:
            on = bit(tq_thread(a))
        elif cls == PT_TT:
            on = bit(tq_thread(a)) & bit(tq_thread(b))
        elif cls == PT_TR:
            on = bit(tq_reg(a)) & bit(tq_thread(b))
        else:
            raise AssertionError(cls)
        # block-class terms must name non-tile qubits, thread/register ones tile qubits
        if cls in (PT_BB,):
            assert a not in tq and (b < 0 or b not in tq)
        if cls in (PT_BT, PT_BR):
            assert b not in tq
        vals.append((on, complex(re, im)))
    for on, z in vals:
        if ph['sign']:
            assert z in (1, -1), z
        f = np.where(on == 1, f * z, f)
    return f
