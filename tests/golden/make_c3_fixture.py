"""Write the fixture of the 1000-qubit chi-64 workload, computed by this repository's MPS
oracle (oracle/mps_ref.py) on the CPU: the checker for the GPU run at full size.

Workload: 1000 qubits, |0..0>, brick-wall depth 40 (20 x 500 + 20 x 499 = 19 980
two-site updates) of exp(-0.1 i H) gates from qsim_amd.workloads.brickwall_gates(1000,
40, seed 42), chi_max 64, cutoff 1e-10, FP64.

Recorded (all gauge-invariant): the bond dimensions and norm_log after every layer,
<Z_i> and <Z_i Z_i+1> of the final state, and the overlaps <phi_k|psi> with four seeded
random product states (log-magnitude and phase, norm_log included) as a fidelity anchor.

usage: python tests/golden/make_c3_fixture.py      (a few minutes, one BLAS thread)
"""

import os
import sys
import time

# 128 x 128 SVDs run fastest on one thread; more threads only contend
os.environ.setdefault('OPENBLAS_NUM_THREADS', '1')
os.environ.setdefault('OMP_NUM_THREADS', '1')

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'fpga-mps-simulator_amd'))

from oracle import mps_ref  # noqa: E402
from qsim_amd import workloads  # noqa: E402

N, DEPTH, CHI, CUTOFF, SEED = 1000, 40, 64, 1e-10, 42
PRODUCT_SEEDS = (1, 2, 3, 4)


def main():
    gates = workloads.brickwall_gates(N, DEPTH, SEED)
    m = mps_ref.MPS(N, CHI, CUTOFF)
    bonds, logs = [], []
    t0 = time.time()
    lay = 0
    for layer, i, g in gates:
        if layer != lay:
            bonds.append(m.bond_dims())
            logs.append(m.norm_log)
            print(f'layer {lay} done, {time.time() - t0:.0f} s, max bond {m.max_bond()}',
                  flush=True)
            lay = layer
        m.apply_adjacent_2q(i, g)
    bonds.append(m.bond_dims())
    logs.append(m.norm_log)
    z = mps_ref.local_expectations(m, mps_ref.FIXED_GATES['z'])
    zz = mps_ref.zz_neighbours(m)
    phis = [mps_ref.random_product_state(N, s) for s in PRODUCT_SEEDS]
    ov = [mps_ref.product_overlap_log(m, p) for p in phis]
    np.savez_compressed(os.path.join(HERE, 'c3_1000q_chi64.npz'),
                        n=N, depth=DEPTH, chi_max=CHI, cutoff=CUTOFF, seed=SEED,
                        product_seeds=np.array(PRODUCT_SEEDS),
                        bond_dims_per_layer=np.array(bonds, dtype=np.int16),
                        norm_log_per_layer=np.array(logs), z=z, zz=zz,
                        overlap_logabs=np.array([o[0] for o in ov]),
                        overlap_phase=np.array([o[1] for o in ov]))
    print(f'done in {time.time() - t0:.0f} s: norm_log {m.norm_log:.12f}, '
          f'mean bond {np.mean(m.bond_dims()):.2f}', flush=True)


if __name__ == '__main__':
    main()
