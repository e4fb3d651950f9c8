"""Compile the engine's C-ABI shared library (liboxr.so) for gfx950 with hipcc.

Every HIP translation unit is built with -ffp-contract=off and correctly rounded fp32
divide/sqrt (the device code's float semantics depend on both) and with packed fp32
instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(PKG), 'csrc')
BUILD = os.path.join(os.path.dirname(PKG), '_build')
LIB = os.path.join(PKG, 'liboxr.so')
SOURCES = ['oxr_kernels.hip', 'oxr_sim_f32.hip', 'oxr_sim_f32_simple.hip', 'oxr_sim_f32_simple_lds.hip', 'oxr_sim_f32_list.hip', 'oxr_sim_f64.hip', 'oxr_sim_f64_list.hip', 'oxr_merge.hip', 'oxr_noalloc.hip', 'oxr_engine.cpp', 'oxr_bvh.cpp', 'oxr_server.cpp', 'oxr_foundry.cpp']
ARCH = os.environ.get('OXR_OFFLOAD_ARCH', 'gfx950')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}', '-ffp-contract=off',
         '-fhip-fp32-correctly-rounded-divide-sqrt',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-command-line-argument']


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src}:\n{r.stderr}')
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, f'--offload-arch={ARCH}', '-shared', '-o', tmp] + objs + ['-lpthread']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{r.stderr}')
    os.replace(tmp, LIB)
    build_examples()
    if verbose:
        print('built', LIB, file=sys.stderr)
    return LIB


EXAMPLES = {'photon_file_source': 'photon_file_source.cpp'}


def build_examples():
    """Host programs over the C ABI and the C++ facade, linked against liboxr.so with an
    rpath to the package directory."""
    inc = os.path.join(os.path.dirname(os.path.dirname(PKG)), 'include')
    src_dir = os.path.join(os.path.dirname(PKG), 'examples')
    for name, src in EXAMPLES.items():
        out = os.path.join(PKG, name)
        cmd = ['g++', '-O2', '-std=c++17', '-Wall', '-I', inc, os.path.join(src_dir, src), '-o', out + '.tmp',
               '-L', PKG, '-loxr', '-Wl,-rpath,$ORIGIN']
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f'g++ failed for {src}:\n{r.stderr}')
        os.replace(out + '.tmp', out)


if __name__ == '__main__':
    build(verbose=True)
