"""Build libgriffon_amd.so (the C-ABI library with the gfx950 HIP kernels) in-tree with hipcc.

Usage: python -m spitfire_amd.build   (or spitfire_amd.build.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIBDIR = os.path.join(HERE, '_lib')
LIB = os.path.join(LIBDIR, 'libgriffon_amd.so')
SOURCES = ['gf_twist_4.hip', 'gf_twist_3.hip', 'gf_twist_2.hip', 'gf_twist_1.hip', 'gf_chem.hip', 'gf_btd.hip',
           'gf_reactor.hip', 'gf_reactor_ext.hip',
           'gf_flamelet.hip', 'gf_eig.hip', 'gf_mech.cpp', 'gf_capi.cpp', 'gf_fbatch.cpp']
ARCH = os.environ.get('GF_OFFLOAD_ARCH', 'gfx950')


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError('hipcc not found')


def _compile(src, obj, verbose):
    cmd = [hipcc(), '--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC', '-c',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',  # no packed FP32 instructions
           '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
           '-I', os.path.join(HERE, '..', 'include'), '-o', obj, src]
    if verbose:
        print(' '.join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose=False, jobs=None):
    """Compile every source to an object (in parallel), then link the shared library."""
    objdir = os.path.join(LIBDIR, 'obj')
    os.makedirs(objdir, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + '.o') for s in SOURCES]
    jobs = jobs or max(1, min(len(srcs), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda so: _compile(so[0], so[1], verbose), zip(srcs, objs)))
    for src, r in zip(SOURCES, results):
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError('hipcc failed compiling ' + src)
    tmp = LIB + '.tmp'
    cmd = [hipcc(), '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed linking libgriffon_amd.so')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
