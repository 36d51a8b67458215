"""The native scheduler (csrc/schedule.cpp), the Fourier multigrid setup
(csrc/amg.cpp) and the C-ABI error state (csrc/capi.cpp) built with
AddressSanitizer and UndefinedBehaviorSanitizer (host code only, g++) and
driven through the C ABI on real operators: every direction's schedule and
every multigrid level is checked (the hierarchy also bit for bit against the
production library's), and operator new is made to fail inside the builds,
which must come back as OBTE_ENOMEM without a leak."""
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


def test_multigrid_setup_under_asan_ubsan(tmp_path):
    if shutil.which('g++') is None:
        pytest.skip('no g++')
    from openbte_amd import fourier as F
    from openbte_amd.geometry import pore_cell_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = pore_cell_2d(30.0, 0.25, 20000, seed=6)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    h0 = F.build_fourier_host(geo, np.diag([150.0, 150.0]), bcs, eff)
    h = F.permute_host(h0, F.morton_order(h0.xy))
    data = tmp_path / 'operator.bin'
    with open(data, 'wb') as f:
        np.array([h.n, int(h.row_ptr[-1])], np.int64).tofile(f)
        np.asarray(h.row_ptr, np.int32).tofile(f)
        np.asarray(h.col, np.int32).tofile(f)
        np.asarray(h.w, np.float64).tofile(f)
        np.asarray(h.diag, np.float64).tofile(f)
    exe = tmp_path / 'amg_sanitize'
    cmd = ['g++', '-O1', '-g', '-std=c++17', '-fsanitize=address,undefined', '-fno-omit-frame-pointer',
           '-fno-sanitize-recover=undefined', '-D__HIP_PLATFORM_AMD__', '-I', os.path.join(ROOT, 'include'),
           '-I', CSRC, '-I', '/opt/rocm/include',
           os.path.join(ROOT, 'tests', 'native', 'amg_sanitize.cpp'), os.path.join(CSRC, 'amg.cpp'),
           os.path.join(CSRC, 'capi.cpp'), '-L/opt/rocm/lib', '-Wl,-rpath,/opt/rocm/lib', '-lamdhip64',
           '-o', str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    out = tmp_path / 'levels.bin'
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=1:verify_asan_link_order=0',
               UBSAN_OPTIONS='print_stacktrace=1:halt_on_error=1',
               LSAN_OPTIONS=f'suppressions={os.path.join(ROOT, "tests", "native", "lsan.supp")}')
    r = subprocess.run([str(exe), str(data), str(out)], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert 'forced failure' in r.stdout
    # the sanitized build exports the production library's hierarchy bit for bit
    hier = F.build_hierarchy(h)
    blob = np.fromfile(out, dtype=np.uint8)
    parts = []
    for l, lv in enumerate(hier.levels):
        nnz_p = int(lv['p_ptr'][-1])
        parts += [lv['p_ptr'], lv['p_col'][:nnz_p], lv['p_val'][:nnz_p]]
        if l > 0:
            nnz_a = int(lv['a_ptr'][-1])
            parts += [lv['a_ptr'], lv['a_col'][:nnz_a], lv['a_w'][:nnz_a], lv['a_diag']]
    expect = np.concatenate([np.ascontiguousarray(p).view(np.uint8) for p in parts])
    n_dense = hier.coarse_inv.shape[0] ** 2 * 8
    assert blob.size == expect.size + n_dense
    assert np.array_equal(blob[:expect.size], expect)
    assert f'ok {len(hier.levels) + 1} levels' in r.stdout
