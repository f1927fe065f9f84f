"""Build recipe for the oracle's C restatement (oracle/sv_kernel_omp.c ->
oracle/libsv_kernel_omp.so, gcc -O3 -fopenmp).  Test infrastructure only.

    python oracle/build_oracle.py
"""

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'sv_kernel_omp.c')
LIB = os.path.join(HERE, 'libsv_kernel_omp.so')


def build(verbose: bool = False) -> str:
    cc = shutil.which('gcc') or 'gcc'
    cmd = [cc, '-O3', '-fopenmp', '-fPIC', '-shared', '-o', LIB + '.tmp', SRC]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + '.tmp', LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
