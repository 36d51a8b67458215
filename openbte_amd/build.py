"""Build libopenbte_amd.so for gfx950 in-tree (hipcc, no JIT cache).

The library holds the HIP kernels (bte_kernels.hip, fourier_kernels.hip, tim_kernels.hip,
material_kernels.hip),
the native host scheduler (schedule.cpp), the Fourier multigrid setup
(amg.cpp) and the C-ABI error state (capi.cpp).  All HIP code is compiled with packed-FP32 instruction
generation disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libopenbte_amd.so')
ARCH = os.environ.get('OBTE_OFFLOAD_ARCH', 'gfx950')

HIP_SOURCES = ['bte_kernels.hip', 'fourier_kernels.hip', 'tim_kernels.hip', 'material_kernels.hip']
CPP_SOURCES = ['schedule.cpp', 'amg.cpp', 'capi.cpp']


def _hipcc() -> str:
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if c and os.path.exists(c):
            return c
    raise RuntimeError('hipcc not found')


def build(verbose: bool = False) -> str:
    hipcc = _hipcc()
    inc = ['-I', os.path.join(ROOT, 'include'), '-I', CSRC]
    with tempfile.TemporaryDirectory(prefix='obte_build_') as tmp:
        objs = []
        jobs = []
        for src in HIP_SOURCES:
            obj = os.path.join(tmp, src + '.o')
            cmd = [hipcc, '-c', '-O3', '-fPIC', '-std=c++17', f'--offload-arch={ARCH}',
                   '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
                   '-ffp-contract=fast'] + inc + [os.path.join(CSRC, src), '-o', obj]
            jobs.append(cmd)
            objs.append(obj)
        for src in CPP_SOURCES:
            obj = os.path.join(tmp, src + '.o')
            cmd = ['g++', '-c', '-O3', '-fPIC', '-std=c++17', '-fopenmp', '-ffp-contract=off',
                   '-I', '/opt/rocm/include', '-D__HIP_PLATFORM_AMD__'] + inc + \
                  [os.path.join(CSRC, src), '-o', obj]
            jobs.append(cmd)
            objs.append(obj)
        procs = [(cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for cmd in jobs]
        for cmd, p in procs:
            out, _ = p.communicate()
            text = out.decode(errors='replace')
            # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
            lines = [ln for ln in text.splitlines() if 'packed-fp32-ops' not in ln]
            if p.returncode != 0:
                raise RuntimeError('compile failed: ' + ' '.join(cmd) + '\n' + '\n'.join(lines))
            if verbose and lines:
                print('\n'.join(lines))
        tmp_lib = LIB + '.tmp'
        cmd = [hipcc, '-shared', '-fPIC', f'--offload-arch={ARCH}', '-fopenmp', '-o', tmp_lib] + objs
        p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if p.returncode != 0:
            raise RuntimeError('link failed: ' + ' '.join(cmd) + '\n' + p.stdout.decode(errors='replace'))
        os.replace(tmp_lib, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
