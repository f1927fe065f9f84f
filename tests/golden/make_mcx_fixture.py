"""Extract the expected basis indices of the reference's MCX stress sweep.

Source: mcx_stress_results.csv (a results file the reference ships, written by
mcx.py:437-575).  Each row names a prepared control pattern and the basis index
the circuit must land on with probability 1.  Only the data columns the tests
need are kept.

usage: python tests/golden/make_mcx_fixture.py [/path/to/reference]
"""
import csv
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'

with open(os.path.join(REF, 'mcx_stress_results.csv'), newline='') as f:
    rows = list(csv.DictReader(f))
with open(os.path.join(HERE, 'mcx_stress_expected.csv'), 'w', newline='') as f:
    w = csv.writer(f)
    w.writerow(['n_qubits', 'test_type', 'pattern_bits', 'expected_idx', 'prob_target'])
    for r in rows:
        w.writerow([r['n_qubits'], r['test_type'], r['pattern_bits'], r['expected_idx'],
                    r['prob_target']])
print(len(rows), 'rows')
