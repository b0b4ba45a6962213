"""Builds libswirlc_amd.so (HIP kernels + C ABI) for gfx950 with hipcc.

Usage: python swirl-c_amd/build_native.py [--jobs N]

Objects go to swirl-c_amd/_build/, the shared library to
swirl-c_amd/swirlc_amd/lib/libswirlc_amd.so (loaded by swirlc_amd._native).
"""

import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIBDIR = os.path.join(HERE, 'swirlc_amd', 'lib')
LIB = os.path.join(LIBDIR, 'libswirlc_amd.so')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('SWIRLC_OFFLOAD_ARCH', 'gfx950')

CFLAGS = [
    f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC',
    # Packed FP32 instructions are kept out of all device code.
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    # Multiply-adds are fused within one expression only, so a function's
    # rounding does not depend on the kernel it is inlined into (the batched
    # and single-grid sweeps agree bit for bit).
    '-ffp-contract=on',
    '-Wno-unused-command-line-argument',
]
# Extra compiler flags (e.g. -D definitions of the kernels' compile-time
# knobs), whitespace separated.
CFLAGS += os.environ.get('SWIRLC_EXTRA_CFLAGS', '').split()


def sources():
  return sorted(f for f in os.listdir(CSRC) if f.endswith('.hip'))


def compile_one(src):
  obj = os.path.join(BUILD, src.replace('.hip', '.o'))
  cmd = [HIPCC, *CFLAGS, '-c', os.path.join(CSRC, src), '-o', obj]
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError(f'hipcc failed for {src}:\n{r.stdout}\n{r.stderr}')
  return obj


def build(jobs=None):
  os.makedirs(BUILD, exist_ok=True)
  os.makedirs(LIBDIR, exist_ok=True)
  jobs = jobs or min(8, os.cpu_count() or 4)
  with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
    objs = list(ex.map(compile_one, sources()))
  tmp = LIB + '.tmp'
  cmd = [HIPCC, f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp, *objs]
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError(f'link failed:\n{r.stdout}\n{r.stderr}')
  shutil.move(tmp, LIB)
  return LIB


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument('--jobs', type=int, default=None)
  args = ap.parse_args()
  print(build(args.jobs))


if __name__ == '__main__':
  sys.exit(main())
