"""The C-ABI library loads and exports every symbol include/openbte_amd.h
declares (no compute calls: no GPU is needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    from openbte_amd import _lib
    lib = _lib.load()
    header = open(os.path.join(ROOT, 'include', 'openbte_amd.h')).read()
    names = sorted(set(re.findall(r'\b(obte_[a-z0-9_]+)\s*\(', header)))
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTED) <= set(names)


def test_error_convention_without_gpu():
    from openbte_amd import _lib
    lib = _lib.load()
    rc = lib.obte_bte_assemble(None, None)
    assert rc < 0
    assert b'plan' in lib.obte_last_error()
    h = ctypes.c_void_p()
    rc = lib.obte_sched_create(0, 0, 2, None, None, None, None, 1, None, 0, ctypes.byref(h))
    assert rc == -1


def test_struct_layout_matches_header():
    """Every field of the C structs is mirrored, in order, by the ctypes structs."""
    from openbte_amd import _lib
    header = open(os.path.join(ROOT, 'include', 'openbte_amd.h')).read()
    for cname, py in (('obte_bte_plan', _lib.BtePlan), ('obte_gcr', _lib.Gcr), ('obte_fourier_plan', _lib.FourierPlan),
                      ('obte_sweep_io', _lib.SweepIO), ('obte_reduce_io', _lib.ReduceIO),
                      ('obte_bte_batch', _lib.Batch), ('obte_amg_level', _lib.AmgLevel)):
        body = re.search(r'typedef struct ' + cname + r' \{(.*?)\}', header, re.S).group(1)
        body = re.sub(r'/\*.*?\*/', '', body, flags=re.S)
        fields = re.findall(r'^\s*(?:const\s+)?\w+\s*\**\s*(\w+)\s*;', body, re.M)
        assert fields == [f[0] for f in py._fields_], cname


def test_sweep_window_rule():
    """The LDS window is the smallest power of two (>= 64) holding two adjacent
    levels, within 48 KB at the chunk's bin count, and only when enabled."""
    from openbte_amd import _lib
    lib = _lib.load()
    assert lib.obte_sweep_window(1, 384, 8) == 512
    assert lib.obte_sweep_window(1, 10, 8) == 64
    assert lib.obte_sweep_window(1, 700, 4) == 1024
    assert lib.obte_sweep_window(1, 513, 8) == 0        # 1024 x 64 B > 48 KB
    assert lib.obte_sweep_window(1, 100, 32) == 0       # wide chunks keep the global path
    assert lib.obte_sweep_window(0, 100, 8) == 0        # disabled
    assert lib.obte_sweep_window(1, 0, 8) == 0          # level sizes unknown


def test_struct_offsets_match_c_compiler(tmp_path):
    """The ctypes mirrors have the C compiler's field offsets and struct sizes
    (types, not only names and order)."""
    import shutil
    import subprocess
    from openbte_amd import _lib
    cc = shutil.which('gcc') or shutil.which('cc')
    if cc is None:
        pytest.skip('no C compiler')
    structs = (('obte_bte_plan_t', _lib.BtePlan), ('obte_gcr_t', _lib.Gcr), ('obte_fourier_plan_t', _lib.FourierPlan),
               ('obte_sweep_io_t', _lib.SweepIO), ('obte_reduce_io_t', _lib.ReduceIO),
               ('obte_bte_batch_t', _lib.Batch), ('obte_amg_level_t', _lib.AmgLevel))
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "openbte_amd.h"', 'int main(void) {']
    for cname, py in structs:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ['  return 0;', '}']
    src = tmp_path / 'offsets.c'
    src.write_text('\n'.join(lines) + '\n')
    exe = tmp_path / 'offsets'
    subprocess.run([cc, '-std=c11', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        s, f, v = line.split()
        got[(s, f)] = int(v)
    for cname, py in structs:
        assert got[(cname, 'size')] == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])


def test_noalloc_fourier_rejects_oversized_plan():
    """The no-allocation Fourier solve refuses plans beyond its fixed
    workspace before touching the device (no GPU needed)."""
    from openbte_amd import _lib
    lib = _lib.load()
    P = _lib.FourierPlan()
    P.n, P.dim, P.nv, P.pin = 262_145, 2, 3, -1
    it, rr = ctypes.c_int32(0), ctypes.c_double(0.0)
    rc = lib.obte_fourier_solve_noalloc(ctypes.byref(P), None, None, None, 1e-13, 10, 16, ctypes.byref(it),
                                        ctypes.byref(rr), None)
    assert rc == -5  # OBTE_ELIMIT
    assert b'OBTE_FOURIER_NOALLOC_MAX_CELLS' in lib.obte_last_error()


def test_host_thread_budget():
    """obte_set_host_threads sets the calling thread's budget for the native
    scheduler's parallel loops: the schedule is the same with 1 and with 4
    threads, and a budget below 1 is rejected."""
    import threading
    import numpy as np
    from openbte_amd import _lib
    from openbte_amd.operator import _build_schedule
    from openbte_amd.geometry import pore_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    from openbte_amd.operator import build_host_operator
    geo = pore_cell_2d(10.0, 0.25, 2000, seed=5)
    mat = synthetic_material_2d(4, 12, 1.0, 1e4)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    op = build_host_operator(geo, mat, bcs, eff)
    out = {}

    def run(k):
        _lib.call('obte_set_host_threads', k)
        out[k] = _build_schedule(op, op.dirs, op.cut_candidate, 192)

    for k in (1, 4):
        t = threading.Thread(target=run, args=(k,))
        t.start()
        t.join()
    for name in ('order', 'lvl_ptr', 'in_src', 'cut_exit', 'nb_src'):
        assert np.array_equal(getattr(out[1], name), getattr(out[4], name))
    with pytest.raises(_lib.ObteError):
        _lib.call('obte_set_host_threads', 0)
