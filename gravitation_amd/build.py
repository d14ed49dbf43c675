"""Builds libgravhip.so (HIP, gfx950) in-tree with hipcc.

Usage: ``python -m gravitation_amd.build [--asm]``.  ``--asm`` also writes the device
assembly of every source next to the build folder (``gravitation_amd/_build/*.s``) for
inspection of the generated gfx950 code.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB_PATH = os.path.join(HERE, 'libgravhip.so')
ARCH = 'gfx950'
SOURCES = ['capi.hip', 'stage1.hip', 'stage1_sym.hip', 'stage2.hip', 'batched.hip',
           'adjoint.hip', 'noalloc.hip']

HIPCC_FLAGS = [
    '-O3', '-std=c++17', '-fPIC', '--offload-arch=%s' % ARCH,
    # Packed FP32 (v_pk_*_f32) is switched off for every device function.
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function',
]


def hipcc():
    rocm = os.environ.get('ROCM_PATH', '/opt/rocm')
    path = os.path.join(rocm, 'bin', 'hipcc')
    return path if os.path.exists(path) else 'hipcc'


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; that is harmless.
    lines = [ln for ln in proc.stdout.splitlines() if 'packed-fp32-ops' not in ln
             and 'not a recognized feature' not in ln]
    if proc.returncode != 0:
        raise RuntimeError('build failed (%d): %s\n%s' % (proc.returncode, ' '.join(cmd),
                                                           '\n'.join(lines)))
    return '\n'.join(lines)


def build(asm=False, verbose=False):
    """Compiles every HIP source into gravitation_amd/libgravhip.so; returns its path."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + '.tmp'
    out = _run([hipcc()] + HIPCC_FLAGS + ['-shared', '-o', tmp] + srcs)
    os.replace(tmp, LIB_PATH)
    if verbose and out.strip():
        print(out)
    if asm:
        bdir = os.path.join(HERE, '_build')
        os.makedirs(bdir, exist_ok=True)
        for s in srcs:
            dst = os.path.join(bdir, os.path.basename(s).replace('.hip', '.s'))
            _run([hipcc()] + HIPCC_FLAGS + ['--cuda-device-only', '-S', '-o', dst, s])
    return LIB_PATH


if __name__ == '__main__':
    print(build(asm='--asm' in sys.argv, verbose=True))
