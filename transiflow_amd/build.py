"""Compile the HIP sources of the backend into ``transiflow_amd/libtransiflow_hip.so``.

Every translation unit is compiled for gfx950 with hipcc and linked into one
shared library that exposes the C ABI declared in ``include/transiflow_hip.h``.
The library is built in-tree so that it travels with the source tree.

Per-file flags:

* assemble.hip -- ``-ffp-contract=off -fno-signed-zeros -ffinite-math-only``:
  products and sums round exactly as the reference's numpy arithmetic (no FMA
  contraction, no reassociation) while constant-zero atom entries fold away;
* every file -- ``-Xclang -target-feature -Xclang -packed-fp32-ops`` keeps the
  compiler from emitting packed FP32 instructions.

``build(force=True)`` recompiles everything; ``force=False`` reuses objects in
``_build/`` whose sources, headers and flags are unchanged.

Usage: ``python -m transiflow_amd.build [-v] [--cached]``
"""

import concurrent.futures
import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libtransiflow_hip.so')
CACHE = os.path.join(os.path.dirname(HERE), '_build')
ARCH = 'gfx950'

COMMON = ['-O3', '--offload-arch=' + ARCH, '-fPIC', '-std=c++17', '-Wall', '-Wno-unused-function',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-I', CSRC, '-I', os.path.join(os.path.dirname(HERE), 'include')]

SOURCES = {
    'assemble.hip': ['-ffp-contract=off', '-fno-signed-zeros', '-ffinite-math-only'],
    'blas.hip': [],
    'krylov.hip': [],
    'linelu.hip': [],
    'bcr.hip': [],
    'gemm.hip': [],
    'conv.hip': [],
    'gmres.hip': [],
    'lineilu.hip': [],
    'mg.hip': [],
}


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc'):
        if cand and os.path.exists(cand):
            return cand
    return 'hipcc'


def _key(src, flags):
    h = hashlib.sha256()
    h.update(' '.join(COMMON + flags).encode())
    for name in sorted(os.listdir(CSRC)):
        if name.endswith('.h') or name == src:
            with open(os.path.join(CSRC, name), 'rb') as f:
                h.update(name.encode() + f.read())
    return h.hexdigest()[:16]


def _compile(cc, src, flags, force, verbose):
    os.makedirs(CACHE, exist_ok=True)
    obj = os.path.join(CACHE, '%s.%s.o' % (src, _key(src, flags)))
    if not force and os.path.exists(obj):
        return obj
    cmd = [cc, '-c', os.path.join(CSRC, src), '-o', obj + '.tmp'] + COMMON + flags
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed on %s:\n%s' % (src, res.stderr))
    os.replace(obj + '.tmp', obj)
    return obj


def build(verbose=False, force=True):
    """Compile all sources and link the shared library; raises on failure."""
    cc = hipcc()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        futs = [ex.submit(_compile, cc, src, flags, force, verbose) for src, flags in SOURCES.items()]
        objs = [f.result() for f in futs]
    tmp_lib = LIB + '.tmp'
    cmd = [cc, '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp_lib] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('link failed:\n%s' % res.stderr)
    os.replace(tmp_lib, LIB)
    return LIB


if __name__ == '__main__':
    build(verbose='-v' in sys.argv, force='--cached' not in sys.argv)
    print(LIB)
