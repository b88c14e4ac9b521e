"""Builds libdiffvg_amd.so for gfx950 with hipcc (in-tree, next to this file)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libdiffvg_amd.so')
SOURCES = ['dvg_host.cpp', 'dvg_kernels.hip']
HEADERS = ['dvg_scene.h', 'dvg_math.h', 'dvg_crmath.h', 'dvg_query.h', 'dvg_winding_fast.h', 'dvg_shade.h',
           'dvg_internal.h', 'dvg_build.h', 'dvg_sort.h']

# -ffp-contract=off keeps every multiply and add separately rounded, as the reference's
# x86-64 CPU build computes them, so coverage decisions are taken on identical bits.
# Packed FP32 instructions are disabled.
HIPCC_FLAGS = ['-O3', '--offload-arch=gfx950', '-std=c++17', '-fPIC', '-ffp-contract=off', '-fno-fast-math',
               '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-munsafe-fp-atomics',
               '-Wno-unused-result']


def build(verbose=False):
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src + '.o')
        lang = ['-x', 'hip'] if src.endswith('.hip') else ['-x', 'hip']
        cmd = [hipcc] + lang + HIPCC_FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + '.tmp'
    cmd = [hipcc, '--offload-arch=gfx950', '-shared', '-fPIC'] + objs + ['-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


DELIVER_SRC = os.path.join(CSRC, 'dvg_deliver.cpp')
DELIVER_LIB = os.path.join(HERE, '_dvg_deliver.so')


def build_deliver(verbose=False):
    """Builds the host-side gradient delivery helper (a torch C++ extension, no device
    code) into _dvg_deliver.so next to this file."""
    import shutil
    import tempfile
    from torch.utils import cpp_extension
    tmp = tempfile.mkdtemp(prefix='dvg_deliver_')
    try:
        mod = cpp_extension.load(name='_dvg_deliver', sources=[DELIVER_SRC], build_directory=tmp,
                                 extra_cflags=['-O2'], verbose=verbose)
        shutil.copyfile(mod.__file__, DELIVER_LIB + '.tmp')
        os.replace(DELIVER_LIB + '.tmp', DELIVER_LIB)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return DELIVER_LIB


if __name__ == '__main__':
    build(verbose='-v' in sys.argv)
    build_deliver(verbose='-v' in sys.argv)
    print(LIB)
    print(DELIVER_LIB)
