"""Builds libmpdo_gfx950.so (the gfx950 HIP kernels + C ABI) in-tree with hipcc.

``python -m mpdo_amd._build`` or ``build()``; objects go to ``csrc/_build/``, the shared
library next to this file so that it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                       # tomography-assisted-mpdo-qcircuit_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, 'csrc')
INCLUDE = os.path.join(REPO, 'include')
LIB = os.path.join(PKG, 'libmpdo_gfx950.so')
SOURCES = ['prof.hip', 'gemm.hip', 'gemm_bf16.hip', 'eigh.hip', 'chol.hip', 'tailcut.hip', 'site_ops.hip', 'layer.hip', 'sample.hip',
           'trace.hip', 'api.hip', 'noalloc.hip']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['-O3', '-std=c++17', '--offload-arch=gfx950', '-fPIC',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-result', '-I', INCLUDE, '-I', CSRC]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError('build failed:\n' + ' '.join(cmd) + '\n' + p.stdout)
    return p.stdout


def build(verbose: bool = False) -> str:
    out = os.path.join(CSRC, '_build')
    os.makedirs(out, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(out, src.replace('.hip', '.o'))
        objs.append(obj)
        cmds.append([HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj])
    with ThreadPoolExecutor(max_workers=min(4, len(cmds))) as ex:
        for log in ex.map(_run, cmds):
            if verbose and log.strip():
                print(log)
    tmp = LIB + '.tmp'
    _run([HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose='-v' in sys.argv)
