"""obte_fourier_solve_noalloc (the entry point that allocates nothing, its
workspace a file-scope device array) against obte_fourier_solve with a
caller workspace, through the C ABI: the same temperatures bit for bit and
the same CG iteration count, on two calls in a row whose second system is
smaller than the first (the array keeps the first call's contents), over the
three solve paths (multigrid multi-launch CG, Jacobi multi-launch CG, one
workgroup)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _device(n_cells, seed, multigrid=True):
    from openbte_amd.fourier import FourierDevice, build_fourier_host
    from openbte_amd.geometry import pore_cell_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = pore_cell_2d(10.0 * np.sqrt(n_cells / 2000.0), 0.25, n_cells, seed=seed)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    host = build_fourier_host(geo, np.diag([150.0, 150.0]), bcs, eff)
    if multigrid:
        return FourierDevice(host)
    return FourierDevice(host, prepared=(host, None, None))


def _solve(dev, noalloc):
    import torch
    from openbte_amd import _lib
    b = dev.b
    b['T'].zero_()
    it, rr = C.c_int32(0), C.c_double(0.0)
    if noalloc:
        _lib.call('obte_fourier_solve_noalloc', dev.plan, b['B'].data_ptr(), b['C'].data_ptr(), b['T'].data_ptr(),
                  1e-13, 20000, 16, C.byref(it), C.byref(rr), dev.stream)
    else:
        _lib.call('obte_fourier_solve', dev.plan, b['B'].data_ptr(), b['C'].data_ptr(), b['T'].data_ptr(),
                  b['ws'].data_ptr(), b['scal'].data_ptr(), 1e-13, 20000, 16, C.byref(it), C.byref(rr), dev.stream)
    torch.cuda.synchronize()
    return b['T'].cpu().numpy().copy(), it.value, rr.value


@pytest.mark.parametrize('sizes', [((20000, True), (2000, True)), ((30000, True), (6000, False))])
def test_noalloc_equals_workspace_solve(sizes):
    for k, (n, mg) in enumerate(sizes):
        dev = _device(n, seed=5 + k, multigrid=mg)
        print(f'# noalloc solve n={dev.h.n} multigrid={dev.perm is not None}', flush=True)
        t_ref, it_ref, rr_ref = _solve(dev, noalloc=False)
        t_na, it_na, rr_na = _solve(dev, noalloc=True)
        print(f'  iterations {it_ref} / {it_na}, relres {rr_ref:.2e} / {rr_na:.2e}, '
              f'max |dT| {np.abs(t_ref - t_na).max():.2e}', flush=True)
        assert it_na == it_ref
        assert rr_ref < 1e-12
        assert np.array_equal(t_na, t_ref)
        del dev
