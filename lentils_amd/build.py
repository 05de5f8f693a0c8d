"""Builds liblentils_hip.so in-tree for gfx950 with hipcc.

Each HIP translation unit is compiled to an object (in parallel), then
linked with rocFFT.  The deflection and rasterization units are compiled
without FMA contraction and with wrapping int arithmetic: their results are
compared bit for bit with the reference C code.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'liblentils_hip.so')
ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')
ARCH = os.environ.get('LENTILS_ARCH', 'gfx950')

COMMON = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}', '-munsafe-fp-atomics',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wno-unused-result',
          f'-I{ROCM}/include']
EXACT = ['-ffp-contract=off', '-fwrapv']
# translation units whose results must match the reference bit for bit
EXACT_UNITS = {'deflect.hip', 'raster.hip', 'triangles.hip'}


def hipcc():
    p = shutil.which('hipcc') or os.path.join(ROCM, 'bin', 'hipcc')
    if not os.path.exists(p):
        raise RuntimeError('hipcc not found')
    return p


def _compile(src, obj):
    flags = COMMON + (EXACT if os.path.basename(src) in EXACT_UNITS else [])
    cmd = [hipcc(), *flags, '-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'compile failed: {" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    return obj


def build(verbose=True):
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith('.hip'))
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    jobs = min(len(srcs), int(os.environ.get('MAX_JOBS', '8')))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, os.path.join(CSRC, s), os.path.join(objdir, s + '.o')) for s in srcs]
        objs = [f.result() for f in futs]
    tmp = OUT + '.tmp'
    cmd = [hipcc(), f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp, *objs,
           f'-L{ROCM}/lib', '-lrocfft', f'-Wl,-rpath,{ROCM}/lib']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed: {" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, OUT)
    if verbose:
        print(f'built {OUT}')
    return OUT


if __name__ == '__main__':
    build()
    sys.exit(0)
