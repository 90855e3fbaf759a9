"""
Build of libframecor_hip.so (hipcc for gfx950) and of the pybind11 module framecor._framecor on its C
ABI (g++), in tree: slipcam_amd/libframecor_hip.so, slipcam_amd/framecor/_framecor*.so.

  python -m slipcam_amd.build          # or __graft_entry__.build()
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libframecor_hip.so')
MODULE_SRC = 'framecor_module.cpp'
MODULE = os.path.join(HERE, 'framecor', '_framecor' + sysconfig.get_config_var('EXT_SUFFIX'))
SOURCES = ['capi.hip', 'batch.hip', 'sat.hip', 'int8_corr.hip', 'int8_wide.hip', 'int8_quad.hip', 'fft_corr.hip', 'fft_fast.hip', 'diff.hip',
           'noalloc.hip']
VERSION = '0.1.0'


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found')


def flags():
    return ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-Wall', '-Wno-unused-result',
            # no packed FP32 instructions (v_pk_*_f32) in the device code
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            f'-DFRAMECOR_VERSION="{VERSION}"', '-I', os.path.join(ROOT, 'include'), '-I', CSRC]


def _compile(src):
    obj = os.path.join(BUILD, src.replace('.hip', '.o'))
    cmd = [hipcc()] + flags() + ['-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'hipcc failed on {src}:\n{r.stderr}')
    # the host pass does not know the device feature above: not a problem
    noise = "'-packed-fp32-ops' is not a recognized feature for this target (ignoring feature)"
    err = '\n'.join(line for line in r.stderr.splitlines() if noise not in line)
    if err.strip():
        print(err, file=sys.stderr)
    return obj


def build(verbose=True):
    """Compile every HIP source for gfx950 and link the shared library (always rebuilds)"""
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as pool:
        objs = list(pool.map(_compile, SOURCES))
    tmp = LIB + '.tmp'
    cmd = [hipcc(), '--offload-arch=gfx950', '-shared', '-fPIC', '-Wl,-soname,libframecor_hip.so', '-o', tmp] + objs \
        + ['-lpthread']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'link failed:\n{r.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print(f'built {LIB}')
    build_module(verbose)
    return LIB


def build_module(verbose=True):
    """The pybind11 module _framecor (host C++, g++) linked to libframecor_hip.so beside the package"""
    import pybind11
    tmp = MODULE + '.tmp'
    cmd = ['g++', '-O2', '-std=c++17', '-shared', '-fPIC', '-fvisibility=hidden', '-Wall',
           '-I', sysconfig.get_paths()['include'], '-I', pybind11.get_include(), '-I', os.path.join(ROOT, 'include'),
           os.path.join(CSRC, MODULE_SRC), '-L', HERE, '-l:libframecor_hip.so', "-Wl,-rpath,$ORIGIN/..", '-o', tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'g++ failed on {MODULE_SRC}:\n{r.stderr}')
    os.replace(tmp, MODULE)
    if verbose:
        print(f'built {MODULE}')
    return MODULE


if __name__ == '__main__':
    if sys.argv[1:] == ['--module']:
        build_module()
    else:
        build()
