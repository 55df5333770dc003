"""Build recipe for libfs2hip.so (HIP, gfx950) and the fs._fs CPython module.

    python -m fs2_amd.build            # build everything in place

libfs2hip.so   hipcc --offload-arch=gfx950 over fs2_amd/csrc/*.hip, *.cpp
fs/_fs.so      g++ over fs2_amd/csrc/py/py_fs.cpp, linked to libfs2hip.so
               with an $ORIGIN rpath

Objects go to fs2_amd/_build/ (git-ignored); the two .so files are written
in place so that they travel with the source tree.
"""
import concurrent.futures as cf
import hashlib
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
INCLUDE = os.path.join(ROOT, 'include')
BUILD = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libfs2hip.so')
IOLIB = os.path.join(HERE, 'libfs2io.so')
HDF5_DIR = os.environ.get('HDF5_DIR', '/opt/conda')
PYEXT = os.path.join(HERE, 'fs', '_fs.so')
STAMP = os.path.join(HERE, 'libfs2hip.so.stamp')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = 'gfx950'

HIP_FLAGS = [
    '-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
    # packed FP32 instructions are kept out of every kernel
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-munsafe-fp-atomics',
    '-I', INCLUDE, '-I', CSRC,
    '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
]

SOURCES = ['fft.hip', 'pm_kernels.hip', 'paint_tiled.hip', 'cola_kernels.hip', 'lpt_kernels.hip',
           'adjoint_kernels.hip', 'capi.cpp', 'diff.cpp', 'batch.cpp', 'host_numerics.cpp', 'fof.cpp', 'probe.hip']


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError('build step failed:\n' + ' '.join(cmd) + '\n' + r.stdout)
    return r.stdout


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + '.o')
    lang = ['-x', 'hip'] if src.endswith('.hip') else ['-x', 'hip']
    _run([HIPCC] + lang + HIP_FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj])
    return obj


def build_lib(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + '.tmp'
    _run([HIPCC, f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_iolib():
    """libfs2io.so (HDF5 output) where an HDF5 C library is present; returns
    its path or None."""
    inc, lib = os.path.join(HDF5_DIR, 'include'), os.path.join(HDF5_DIR, 'lib')
    if not (os.path.exists(os.path.join(inc, 'hdf5.h')) and
            os.path.exists(os.path.join(lib, 'libhdf5.so'))):
        return None
    os.makedirs(BUILD, exist_ok=True)
    obj = os.path.join(BUILD, 'io_hdf5.cpp.o')
    _run([HIPCC, '-x', 'hip'] + HIP_FLAGS + ['-I', inc, '-c', os.path.join(CSRC, 'io_hdf5.cpp'),
                                             '-o', obj])
    tmp = IOLIB + '.tmp'
    _run([HIPCC, f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp, obj, '-L', HERE,
          '-lfs2hip', '-L', lib, '-lhdf5', "-Wl,-rpath,$ORIGIN", '-Wl,-rpath,' + lib])
    os.replace(tmp, IOLIB)
    return IOLIB


def build_pyext():
    import numpy
    py_inc = sysconfig.get_paths()['include']
    src = os.path.join(CSRC, 'py', 'py_fs.cpp')
    tmp = PYEXT + '.tmp'
    _run(['g++', '-O2', '-std=c++17', '-fPIC', '-shared', '-Wall',
          '-I', py_inc, '-I', numpy.get_include(), '-I', INCLUDE,
          src, '-o', tmp, '-L', HERE, '-lfs2hip', '-ldl', "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, PYEXT)
    return PYEXT


def source_hash():
    """SHA-256 over every file under csrc/ and include/, this recipe, the
    compiler flags and the HDF5 location: what the built libraries depend on."""
    h = hashlib.sha256()
    h.update(repr((HIP_FLAGS, SOURCES, HDF5_DIR, ARCH)).encode())
    files = [os.path.abspath(__file__)]
    for top in (CSRC, INCLUDE):
        for d, _, names in os.walk(top):
            files += [os.path.join(d, n) for n in names]
    for f in sorted(files):
        h.update(os.path.relpath(f, ROOT).encode())
        with open(f, 'rb') as fp:
            h.update(fp.read())
    return h.hexdigest()


def build_all(jobs=8):
    digest = source_hash()
    build_lib(jobs)
    build_iolib()
    build_pyext()
    with open(STAMP, 'w') as f:
        f.write(digest + '\n')


def is_current():
    """True when both libraries exist and were built from the sources as
    they are now (the stamp next to them holds the same source hash)."""
    if not (os.path.exists(LIB) and os.path.exists(PYEXT) and os.path.exists(STAMP)):
        return False
    with open(STAMP) as f:
        return f.read().strip() == source_hash()


def ensure_built(jobs=8):
    """Builds when the libraries are missing or were built from other
    sources or flags; a current build is loaded as it is."""
    if not is_current():
        print('fs2_amd: building libfs2hip.so and fs._fs (sources changed or no build)',
              file=sys.stderr, flush=True)
        build_all(jobs)


if __name__ == '__main__':
    build_all()
    print('built', LIB, PYEXT)
    sys.exit(0)
