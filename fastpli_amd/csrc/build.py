"""Build libfastpli_amd.so (HIP, gfx950) in-tree with hipcc.

    python -m fastpli_amd.csrc.build        # or fastpli_amd.csrc.build.build()

Every translation unit is compiled for gfx950 only.  Packed FP32 instructions
are disabled for all device code.  The voxelizer is compiled with
-ffp-contract=off: its distance test must round every FP64 operation exactly
once, as the reference's scalar code does.
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT_DIR = os.path.join(PKG, 'lib')
OUT = os.path.join(OUT_DIR, 'libfastpli_amd.so')
BUILD_DIR = os.path.join(PKG, '_build')

ARCH = os.environ.get('FASTPLI_AMD_ARCH', 'gfx950')

COMMON = [
    f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC',
    '-fhip-fp32-correctly-rounded-divide-sqrt', '-Xclang', '-target-feature',
    '-Xclang', '-packed-fp32-ops', '-Wno-unused-result',
    f'-I{os.path.join(ROOT, "include")}'
]

SOURCES = {
    'fpl_api.hip': [],
    'fpl_generate.hip': ['-ffp-contract=off'],
    'fpl_analysis.hip': ['-ffp-contract=off'],
    # tolerance-based FP32 chain: hardware reciprocal / rsqrt instead of the
    # correctly rounded division and square root sequences
    'fpl_simulate.hip': ['-fno-hip-fp32-correctly-rounded-divide-sqrt'],
    'fpl_serve.hip': [],
    'fpl_rofl.hip': ['-ffp-contract=off'],
}


def hipcc():
    path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(path):
        raise RuntimeError('hipcc not found')
    return path


def _compile(src, extra):
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + '.o')
    cmd = [hipcc()] + COMMON + extra + ['-c', os.path.join(HERE, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'compiling {src} failed:\n{" ".join(cmd)}\n{r.stderr}')
    return obj


def build(verbose=False):
    """Compile all sources and link the shared library; returns its path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    os.makedirs(OUT_DIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        futs = [ex.submit(_compile, s, e) for s, e in SOURCES.items()]
        objs = [f.result() for f in futs]
    tmp = OUT + '.tmp'
    cmd = [hipcc(), f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'linking failed:\n{" ".join(cmd)}\n{r.stderr}')
    os.replace(tmp, OUT)
    if verbose:
        print('built', OUT)
    return OUT


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
