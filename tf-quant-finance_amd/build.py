"""Builds libtqf.so (all HIP sources, gfx950) in-tree.

  python tf-quant-finance_amd/build.py [--jobs N]

Each source compiles to an object under tqf/_lib/obj/ in parallel, then one
shared library links them: tqf/_lib/libtqf.so.  No JIT cache is used, so the
library travels with the tree.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIBDIR = os.path.join(HERE, 'tqf', '_lib')
OBJDIR = os.path.join(LIBDIR, 'obj')
LIB = os.path.join(LIBDIR, 'libtqf.so')
ARCH = 'gfx950'

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = [
    '-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH,
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result', '-I' + os.path.join(HERE, '..', 'include'),
]


def sources():
  return sorted(f for f in os.listdir(CSRC)
                if f.endswith('.hip') or f.endswith('.cpp'))


def _compile(src):
  obj = os.path.join(OBJDIR, src + '.o')
  cmd = [HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
  if src.endswith('.cpp'):
    cmd = [HIPCC] + FLAGS + ['-x', 'hip', '-c', os.path.join(CSRC, src), '-o',
                             obj]
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError('compile failed: %s\n%s\n%s' % (' '.join(cmd), r.stdout,
                                                      r.stderr))
  return obj


def build(jobs=None, verbose=True):
  os.makedirs(OBJDIR, exist_ok=True)
  srcs = sources()
  jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
  with cf.ThreadPoolExecutor(jobs) as ex:
    objs = list(ex.map(_compile, srcs))
  tmp = LIB + '.tmp'
  cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError('link failed:\n%s\n%s' % (r.stdout, r.stderr))
  os.replace(tmp, LIB)
  if verbose:
    print('built', LIB)
  return LIB


if __name__ == '__main__':
  ap = argparse.ArgumentParser()
  ap.add_argument('--jobs', type=int, default=None)
  args = ap.parse_args()
  try:
    build(args.jobs)
  except RuntimeError as e:
    print(e, file=sys.stderr)
    sys.exit(1)
