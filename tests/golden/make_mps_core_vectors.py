"""Write golden vectors from the reference's mps_core.py (run in the dev container only).

The bond-dimension trajectory and norm_log along truncating runs depend on
LAPACK rounding at the cutoff/chi_max boundary, which no closed form pins, so
they are recorded from the reference itself.  The reference source is only
imported here; the fixture holds inputs and outputs (gates, bond dims,
norm_log, singular values, <Z_i>, a small statevector), never code.

usage: python tests/golden/make_mps_core_vectors.py [/path/to/reference]
"""

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'


def haar(rng):
    z = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))[None, :]


def brickwall(n, depth):
    return [list(range(layer % 2, n - 1, 2)) for layer in range(depth)]


def run_case(mps_core, n, depth, chi_max, cutoff, seed):
    rng = np.random.default_rng(seed)
    st = mps_core.MPSState(n, chi_max, cutoff)
    gates, sites, keeps, logs = [], [], [], []
    for bonds in brickwall(n, depth):
        for i in bonds:
            g = haar(rng)
            st.apply_2q(i, i + 1, g)
            gates.append(g)
            sites.append(i)
            keeps.append(st.chi_r(i))
            logs.append(st.norm_log)
    out = dict(n=n, depth=depth, chi_max=chi_max, cutoff=cutoff, seed=seed,
               gates=np.array(gates), sites=np.array(sites), keeps=np.array(keeps),
               norm_log_traj=np.array(logs), norm_log=st.norm_log,
               bond_dims=np.array(st.bond_dims()))
    spectra = np.zeros((n - 1, chi_max))
    for b in range(n - 1):
        s = st.bond_spectrum(b)
        spectra[b, :s.size] = s
    out['bond_spectra'] = spectra
    if n <= 16:
        sv = st.get_statevector()
        idx = np.arange(sv.size)
        p = np.abs(sv) ** 2
        out['z_expect'] = np.array([np.sum(p * (1 - 2 * ((idx >> q) & 1))) / np.sum(p)
                                    for q in range(n)])
        out['norm_sq'] = float(np.sum(p))
        if n <= 12:
            out['statevector'] = sv
    return out


def main():
    sys.path.insert(0, REF)
    import mps_core  # noqa: E402  (reference, dev container only)

    cases = {
        # 16 qubits, brick-wall depth 8 of Haar gates, chi_max 16, cutoff 1e-12
        'c1_16q': (16, 8, 16, 1e-12, 2024),
        # 12 qubits, depth 10, chi_max 8: truncates, records the statevector
        'trunc_12q': (12, 10, 8, 1e-12, 7),
        # 10 qubits, depth 6, chi_max 32: no truncation by chi
        'exact_10q': (10, 6, 32, 1e-12, 11),
    }
    for name, args in cases.items():
        data = run_case(mps_core, *args)
        path = os.path.join(HERE, f'mps_core_{name}.npz')
        np.savez_compressed(path, **data)
        print(f'{path}: bond_dims={list(data["bond_dims"])} norm_log={data["norm_log"]:.6e}')


if __name__ == '__main__':
    main()
