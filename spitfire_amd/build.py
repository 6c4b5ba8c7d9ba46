"""Build libgriffon_amd.so (the C-ABI library with the gfx950 HIP kernels) in-tree with hipcc.

Usage: python -m spitfire_amd.build   (or spitfire_amd.build.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIBDIR = os.path.join(HERE, '_lib')
LIB = os.path.join(LIBDIR, 'libgriffon_amd.so')
SOURCES = ['gf_mech.cpp', 'gf_capi.cpp', 'gf_chem.hip', 'gf_btd.hip', 'gf_flamelet.hip', 'gf_eig.hip', 'gf_fbatch.cpp']
ARCH = os.environ.get('GF_OFFLOAD_ARCH', 'gfx950')


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError('hipcc not found')


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + '.tmp'
    cmd = [hipcc(), '--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC', '-shared',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',  # no packed FP32 instructions
           '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
           '-I', os.path.join(HERE, '..', 'include'),
           '-o', tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed building libgriffon_amd.so')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
