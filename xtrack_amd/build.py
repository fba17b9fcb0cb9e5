"""Build libxtrack_amd.so in the package directory (gfx950, in-tree).

    python -m xtrack_amd.build

The kernel is compiled with -ffp-contract=off (no a*b+c fusion beyond the
explicit fma calls of the math library) and with packed FP32 disabled; the
host-side program builder is plain C++ compiled with the same contraction
setting so host-evaluated constants match device arithmetic bit for bit.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, 'csrc')
BUILD = os.path.join(PKG, '_build')
ROOT = os.path.dirname(PKG)

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('XTM_OFFLOAD_ARCH', 'gfx950')

HIP_FLAGS = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}', '-ffp-contract=off',
             '-fno-fast-math', '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-Wno-unused-command-line-argument', f'-I{ROOT}/include', f'-I{CSRC}']
CXX_FLAGS = ['-O2', '-std=c++17', '-fPIC', '-ffp-contract=off', '-fno-fast-math',
             f'-I{ROOT}/include', f'-I{CSRC}']


def _run(cmd):
    print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True, out=None, extra_hip_flags=(), build_dir=None):
    BUILD_ = build_dir or BUILD
    os.makedirs(BUILD_, exist_ok=True)
    kern_o = os.path.join(BUILD_, 'xt_kernel.o')
    bld_o = os.path.join(BUILD_, 'xt_builder.o')
    flo_o = os.path.join(BUILD_, 'xt_flops.o')
    _run([HIPCC, *HIP_FLAGS, *extra_hip_flags, '-c', os.path.join(CSRC, 'xt_kernel.hip'),
          '-o', kern_o])
    _run([shutil.which('g++') or 'g++', *CXX_FLAGS, '-c',
          os.path.join(CSRC, 'xt_builder.cpp'), '-o', bld_o])
    _run([shutil.which('g++') or 'g++', *CXX_FLAGS, '-c',
          os.path.join(CSRC, 'xt_flops.cpp'), '-o', flo_o])
    out = out or os.path.join(PKG, 'libxtrack_amd.so')
    tmp = out + '.tmp'
    _run([HIPCC, '-shared', f'--offload-arch={ARCH}', '-fPIC', '-o', tmp, kern_o, bld_o, flo_o])
    os.replace(tmp, out)
    return out


if __name__ == '__main__':
    build()
    sys.exit(0)
