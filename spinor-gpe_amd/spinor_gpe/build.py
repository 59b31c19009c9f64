"""Build libsgpe.so (HIP kernels + C ABI) in-tree for gfx950.

Usage: ``python -m spinor_gpe.build`` (or ``build()``). Objects go to
``<package root>/_build``; the library to ``spinor_gpe/_lib/libsgpe.so``.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(ROOT, 'csrc')
BUILD = os.path.join(ROOT, '_build')
OUT = os.path.join(PKG, '_lib', 'libsgpe.so')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('SGPE_ARCH', 'gfx950')
# Packed FP32 instructions are disabled: some of their forms have produced
# silently wrong results on this hardware. Kernel arguments are preloaded
# into scalar registers at wave launch (gfx950 supports it; the code object
# keeps a fallback entry that loads them from memory), which takes one
# dependent memory round trip off the start of every kernel.
FLAGS = ['-O3', '-std=c++17', f'--offload-arch={ARCH}', '-fPIC', '-Wno-unused-result',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-mllvm', '-amdgpu-kernarg-preload-count=16']
SOURCES = ['sgpe_api.hip', 'sgpe_inst_f64.hip', 'sgpe_inst_f32.hip', 'sgpe_persist_f64.hip',
           'sgpe_persist_f32.hip', 'sgpe_unwrap.cpp']


def _run(cmd, verbose):
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n"
                           f"{res.stdout}\n{res.stderr}")
    return res


def build(verbose=False):
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    objs = [os.path.join(BUILD, os.path.splitext(s)[0] + '.o') for s in SOURCES]
    cmds = [[HIPCC, *FLAGS, '-c', os.path.join(CSRC, s), '-o', o] for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as pool:
        list(pool.map(lambda c: _run(c, verbose), cmds))
    tmp = OUT + '.tmp'
    _run([HIPCC, *FLAGS, '-shared', '-o', tmp, *objs], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
