"""Replay a fused-pass plan (qs_sv_plan_json) with NumPy, as the pass kernel reads it.

Test helper: it resolves every register bit through the stage's register list
and the pass's tile qubit list, exactly like sv_pass_kernel, and applies the
gates to a dense state.  A planner bug (wrong order, wrong register index,
target outside the tile) shows up as a mismatch against the oracle.
"""

import numpy as np

DG_U1, DG_U2, DG_DIAG1, DG_DIAG2, DG_SIGN2 = 0, 1, 2, 3, 4


def _cm(vals, n):
    v = np.asarray(vals[:2 * n], dtype=np.float64)
    return v[0::2] + 1j * v[1::2]


def _ctrl_ok(idx, cmask):
    return (idx & cmask) == cmask


def replay(plan: dict, nq: int, state: np.ndarray) -> np.ndarray:
    st = state.copy()
    idx = np.arange(st.size, dtype=np.int64)
    for p in plan['passes']:
        tq = p['tq']
        assert len(tq) == p['K'] and sorted(tq) == tq
        tile = 0
        for q in tq:
            tile |= 1 << q
        for s in p['stages']:
            regs = s['regs']
            assert len(regs) == p['R'] and sorted(regs) == regs
            for g in s['gates']:
                cm = int(g['cmask'])
                sel = _ctrl_ok(idx, cm)
                if g['kind'] == DG_U1:
                    t = tq[regs[g['r0']]]
                    m = _cm(g['m'], 4).reshape(2, 2)
                    lo = sel & (((idx >> t) & 1) == 0)
                    i0 = idx[lo]
                    i1 = i0 | (1 << t)
                    a0, a1 = st[i0].copy(), st[i1].copy()
                    st[i0] = m[0, 0] * a0 + m[0, 1] * a1
                    st[i1] = m[1, 0] * a0 + m[1, 1] * a1
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
                elif g['kind'] == DG_SIGN2:
                    hit = sel & (((idx >> g['qa']) & (idx >> g['qb']) & 1) == 1)
                    st[hit] *= -1
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
                if g['kind'] in (DG_U1, DG_U2):
                    assert (tile >> tq[regs[g['r0']]]) & 1
    return st
