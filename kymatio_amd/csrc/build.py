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

# (source, extra defines, object name); compiled in parallel, then linked
PRODUCT_UNITS = [('kpass2d.hip', [], 'kpass2d.o'), ('kspec2d.hip', [], 'kspec2d.o'), ('capi.hip', [], 'capi.o'),
                 ('kinst1d.hip', [], 'kinst1d.o'),
                 ('ops.hip', [], 'ops.o'), ('decoder.hip', [], 'decoder.o')] + [
    ('kinst2d.hip', ['-DKYM_T=%s' % T, '-DKYM_P=%d' % P, '-DKYM_PICK=kym_pick2d_%s_p%d' % (tag, P)],
     'kinst2d_%s_p%d.o' % (tag, P))
    for tag, T in (('f32', 'float'), ('f64', 'double')) for P in (7, 17, 31)]
EMU_SOURCES = ['emu.hip']
BUILD_DIR = os.path.join(HERE, '_build')


def _run(cmd):
    t = time.time()
    print(' '.join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=HERE)
    print('  %.1f s' % (time.time() - t), flush=True)


def build_product(jobs=None):
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(BUILD_DIR, exist_ok=True)
    jobs = jobs or min(len(PRODUCT_UNITS), max(1, os.cpu_count() or 1))
    objs = []

    def compile_unit(u):
        src, defs, obj = u
        o = os.path.join(BUILD_DIR, obj)
        _run([HIPCC, '--offload-arch=' + ARCH, '-c'] + [f for f in COMMON if f != '-shared'] + defs + [src, '-o', o])
        return o

    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_unit, PRODUCT_UNITS))
    out = os.path.join(PKG, 'libkymatio_hip.so')
    tmp = out + '.tmp'
    _run([HIPCC, '--offload-arch=' + ARCH, '-shared', '-fPIC'] + objs + ['-o', tmp])
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
