"""Build recipe for libqsim_hip.so (hipcc, gfx950 only).

    python -m qsim_amd.build          (from fpga-mps-simulator_amd/)

The shared library is written next to this file, in-tree, so that it travels
with the source tree to the machine that runs it.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(ROOT_DIR, 'csrc')
LIB_NAME = 'libqsim_hip.so'
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)

SOURCES = ['abi.hip', 'sv_kernels.hip', 'sv_plan.cpp', 'sv_jit.cpp', 'sv_grad.hip',
           'sv_sample.hip', 'mps_kernels.hip', 'mps_sample.hip', 'mps_readout.hip',
           'mps_grad.hip']

# Headers the per-plan pass kernels are compiled from at run time (sv_jit.cpp): their
# text is embedded into the library as string constants.
JIT_HEADERS = [('kSvDeviceTypesSrc', 'sv_device_types.h'), ('kSvOpsSrc', 'sv_ops.h')]
JIT_INC = 'sv_jit_src.inc'

HIPCC_FLAGS = [
    '--offload-arch=gfx950',
    '-O3', '-std=c++17', '-fPIC', '-shared',
    '-Wall', '-Wno-unused-function',
    # Packed FP32 ops (v_pk_*_f32) are kept out of the generated code.
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
]
LINK_LIBS = ['-lhiprtc']


def hipcc() -> str:
    path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(path):
        raise RuntimeError('hipcc not found: the HIP library cannot be built')
    return path


def write_jit_sources() -> str:
    """Writes csrc/sv_jit_src.inc: each run-time compiled header as a raw string constant."""
    out = []
    for var, name in JIT_HEADERS:
        with open(os.path.join(CSRC, name)) as f:
            text = f.read()
        if ')QSRC"' in text:
            raise RuntimeError(f'{name} contains the raw-string delimiter')
        out.append(f'// text of {name} (written by qsim_amd/build.py)\n'
                   f'const char {var}[] = R"QSRC({text})QSRC";\n')
    path = os.path.join(CSRC, JIT_INC)
    new = '\n'.join(out)
    old = None
    if os.path.exists(path):
        with open(path) as f:
            old = f.read()
    if old != new:
        with open(path, 'w') as f:
            f.write(new)
    return path


def build(verbose: bool = False) -> str:
    write_jit_sources()
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB_PATH + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-o', tmp] + srcs + LINK_LIBS
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f'hipcc failed with exit code {res.returncode}')
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == '__main__':
    print(build(verbose=True))
