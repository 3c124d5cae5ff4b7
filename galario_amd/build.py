"""Build the HIP library in-tree: galario_amd/libgalario_amd.so (gfx950).

    python galario_amd/build.py

hipcc cross-compiles for gfx950 without a GPU present.  Flags:
  -ffp-contract=off     the stages that restate galario's arithmetic round as
                        the reference's x86-64 build (no FMA contraction);
                        the FFT uses explicit fma where it wants one.
  -Xclang -target-feature -Xclang -packed-fp32-ops
                        no packed FP32 instructions are generated.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, 'csrc', 'galario_amd.hip')
OUT = os.path.join(HERE, 'libgalario_amd.so')
ARCH = os.environ.get('GALARIO_OFFLOAD_ARCH', 'gfx950')


def hipcc() -> str:
    for cand in (shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found: the galario_amd HIP library cannot be built')


def command(out: str = OUT) -> list:
    return [hipcc(), f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC', '-shared',
            '-ffp-contract=off', '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-Wall', '-Wno-unused-function', '-Wno-unused-command-line-argument',
            '-I', os.path.join(ROOT, 'include'), '-I', os.path.join(HERE, 'csrc'),
            SRC, '-o', out]


def build(verbose: bool = False) -> str:
    tmp = OUT + '.tmp'
    cmd = command(tmp)
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError('hipcc failed building galario_amd')
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose=True))
