"""The native scheduler (csrc/schedule.cpp) and the C-ABI error state
(csrc/capi.cpp) built with AddressSanitizer and UndefinedBehaviorSanitizer
(host code only, g++) and driven through the C ABI on a real operator: every
direction's schedule is checked, and operator new is made to fail inside the
OpenMP region, which must come back as OBTE_ENOMEM without a leak."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, 'openbte_amd', 'csrc')


def test_scheduler_under_asan_ubsan(tmp_path):
    if shutil.which('g++') is None:
        pytest.skip('no g++')
    from openbte_amd.geometry import pore_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    from openbte_amd.operator import build_host_operator
    geo = pore_cell_2d(10.0, 0.25, 1500, seed=4)
    mat = synthetic_material_2d(4, 12, 1.0, 1e3)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    op = build_host_operator(geo, mat, bcs, EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0))
    data = tmp_path / 'faces.bin'
    with open(data, 'wb') as f:
        np.array([op.n, len(op.i_mat), op.dim, op.n_dirs], np.int64).tofile(f)
        np.asarray(op.i_mat, np.int32).tofile(f)
        np.asarray(op.j_mat, np.int32).tofile(f)
        np.ascontiguousarray(op.k_face, np.float64).tofile(f)
        np.asarray(op.cut_candidate, np.uint8).tofile(f)
        np.ascontiguousarray(op.dirs, np.float64).tofile(f)
    exe = tmp_path / 'sched_sanitize'
    cmd = ['g++', '-O1', '-g', '-std=c++17', '-fopenmp', '-fsanitize=address,undefined', '-fno-omit-frame-pointer',
           '-fno-sanitize-recover=undefined', '-D__HIP_PLATFORM_AMD__', '-I', os.path.join(ROOT, 'include'),
           '-I', CSRC, '-I', '/opt/rocm/include',
           os.path.join(ROOT, 'tests', 'native', 'sched_sanitize.cpp'), os.path.join(CSRC, 'schedule.cpp'),
           os.path.join(CSRC, 'capi.cpp'), '-L/opt/rocm/lib', '-Wl,-rpath,/opt/rocm/lib', '-lamdhip64',
           '-o', str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=1:verify_asan_link_order=0', OMP_NUM_THREADS='4',
               UBSAN_OPTIONS='print_stacktrace=1:halt_on_error=1',
               LSAN_OPTIONS=f'suppressions={os.path.join(ROOT, "tests", "native", "lsan.supp")}')
    r = subprocess.run([str(exe), str(data)], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert 'ok 12 directions' in r.stdout
    assert 'forced failure' in r.stdout
