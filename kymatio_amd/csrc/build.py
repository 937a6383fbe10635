"""Build the HIP libraries in-tree (gfx950).

    python kymatio_amd/csrc/build.py [--emu-only] [--hip-only]

Outputs
    kymatio_amd/libkymatio_hip.so   product library (C ABI in include/kymatio_hip.h)
    kymatio_amd/libkymatio_emu.so   host-only test harness running the same
                                    workgroup bodies on the CPU (tests only)
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ARCH = os.environ.get('KYM_OFFLOAD_ARCH', 'gfx950')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

# -packed-fp32-ops: never emit v_pk_{add,mul,fma}_f32 (see the package notes)
COMMON = ['-O3', '-std=c++17', '-fPIC', '-shared', '-Wno-unused-result',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']

PRODUCT_SOURCES = ['capi.hip']
EMU_SOURCES = ['emu.hip']


def _run(cmd):
    t = time.time()
    print(' '.join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=HERE)
    print('  %.1f s' % (time.time() - t), flush=True)


def build_product():
    out = os.path.join(PKG, 'libkymatio_hip.so')
    tmp = out + '.tmp'
    _run([HIPCC, '--offload-arch=' + ARCH] + COMMON + PRODUCT_SOURCES + ['-o', tmp])
    os.replace(tmp, out)
    return out


def build_emu():
    out = os.path.join(PKG, 'libkymatio_emu.so')
    tmp = out + '.tmp'
    _run([HIPCC, '-x', 'hip', '--offload-host-only', '-O2', '-std=c++17', '-fPIC', '-shared']
         + EMU_SOURCES + ['-o', tmp])
    os.replace(tmp, out)
    return out


def main(argv):
    if '--hip-only' not in argv:
        build_emu()
    if '--emu-only' not in argv:
        build_product()


if __name__ == '__main__':
    main(sys.argv[1:])
