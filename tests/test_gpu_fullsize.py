"""Properties at the full 200k-cell size (199,713 triangles, 60 MFP bins x 96
directions), where the oracle is too slow to run: mirror symmetry of the
discrete problem and agreement of the two sweep variants."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case():
    from openbte_amd.geometry import pore_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = pore_cell_2d(100.0, 0.25, 200_000, seed=0)
    mat = synthetic_material_2d(60, 96, 1.0, 1.0e5, seed=0)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    return geo, mat, bcs, eff


def _iterate(geo, mat, bcs, eff, n_iter, window=True):
    import torch
    from openbte_amd.bte import BTEDevice, SweepConfig
    from openbte_amd.fourier import Fourier
    from openbte_amd.operator import build_host_operator
    t0 = Fourier(geo, mat.thermal_conductivity, bcs, eff, verbose=False).variables['Temperature_Fourier']['data']
    dev = BTEDevice(build_host_operator(geo, mat, bcs, eff), cfg=SweepConfig(window=window))
    dev.reset(t0)
    for _ in range(n_iter):
        dev.outer_iteration()
    torch.cuda.synchronize()
    out = (float(dev.buf['kappa'].sum().item()), dev.buf['dt'][dev.cur].cpu().numpy().copy())
    del dev
    torch.cuda.empty_cache()
    return out


def test_mirror_symmetry_full_size(tmp_path):
    """Reflecting the cell in y maps direction phi to pi - phi (in the set for
    an even direction count) and leaves the x-contact unchanged: kappa and the
    temperature of mirrored cells agree up to the interface tolerance.  The
    mirrored mesh goes through the gmsh writer and reader."""
    from scipy.spatial import cKDTree
    from openbte_amd.msh import read_msh, write_msh
    geo, mat, bcs, eff = _case()
    assert geo.n_elems > 190_000
    fn = str(tmp_path / 'cell.msh')
    write_msh(geo, fn)
    lines = open(fn).read().split('\n')
    i0 = lines.index('$Nodes') + 2
    n = int(lines[i0 - 1])
    for i in range(i0, i0 + n):
        k, x, y, z = lines[i].split()
        lines[i] = f'{k} {x} {-float(y)!r} {z}'
    fm = str(tmp_path / 'mirror.msh')
    open(fm, 'w').write('\n'.join(lines))
    geo_m = read_msh(fm)
    assert geo_m.n_elems == geo.n_elems
    print('# full size mirror', geo.n_elems, flush=True)
    k, dt = _iterate(geo, mat, bcs, eff, 3)
    km, dtm = _iterate(geo_m, mat, bcs, eff, 3)
    c = np.asarray(geo.elem_centroids)
    cm = np.asarray(geo_m.elem_centroids) * np.array([1.0, -1.0])
    dist, idx = cKDTree(c).query(cm)
    assert dist.max() < 1e-9 * 100.0
    print('kappa', k, km, 'rel', abs(km / k - 1), 'dT', np.abs(dtm - dt[idx]).max())
    assert abs(km / k - 1) < 1e-8
    assert np.abs(dtm - dt[idx]).max() < 1e-8


def test_window_equals_global_full_size():
    """The LDS-window sweep and the all-HBM sweep at full size: the same kappa
    and temperatures bit for bit after two outer iterations."""
    geo, mat, bcs, eff = _case()
    print('# full size window', geo.n_elems, flush=True)
    k1, dt1 = _iterate(geo, mat, bcs, eff, 2, window=True)
    k2, dt2 = _iterate(geo, mat, bcs, eff, 2, window=False)
    assert k1 == k2
    assert np.array_equal(dt1, dt2)
