"""Build libemc_hip.so (the HIP engine + C ABI) in-tree with hipcc for gfx950.

Usage: ``python -m dragonfly_amd.build`` (also called by ``__graft_entry__.build()``).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libemc_hip.so')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')

SOURCES = ['emc_engine.hip', 'emc_kernels.hip', 'emc_mp.hip', 'emc_merge.hip', 'emc_nn.hip', 'emc_subset.hip', 'emc_background.hip',
           'quat_gen.cpp']

# -ffp-contract=off: no fused multiply-add anywhere, so every FP64 value keeps the
# reference's rounding sequence. Packed FP32 ops are disabled for device code.
FLAGS = ['-O3', '-std=c++17', '--offload-arch=gfx950', '-fPIC', '-shared', '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wall',
         '-Wno-unused-function', '-I' + INCLUDE]


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError('hipcc not found')


def build(verbose=False):
    tmp = LIB + '.tmp'
    cmd = [hipcc()] + FLAGS + ['-o', tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError('hipcc failed building libemc_hip.so')
    # every symbol must resolve now, not at the first call on a GPU machine
    nm = subprocess.run(['nm', '-D', '--undefined-only', tmp], capture_output=True, text=True)
    missing = [l.split()[-1] for l in nm.stdout.splitlines() if '_ZN3emc' in l]
    if missing:
        os.remove(tmp)
        raise RuntimeError('libemc_hip.so has unresolved engine symbols: %s' % ', '.join(missing))
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
