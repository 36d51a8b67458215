"""Results container: state files round-trip without pickle, VTK output
layout, node averaging and tiling."""
import numpy as np

from openbte_amd.geometry import pore_cell_2d, pore_slab_3d
from openbte_amd.material import synthetic_material_2d
from openbte_amd.objects import SolverResults
from openbte_amd.results import OpenBTEResults


def _results(geo):
    rng = np.random.default_rng(0)
    T = rng.standard_normal(geo.n_elems)
    J = rng.standard_normal((geo.n_elems, geo.dim))
    bte = SolverResults(kappa_eff=12.5, variables={'Temperature_BTE': {'data': T, 'units': 'K'},
                                                   'Flux_BTE': {'data': J, 'units': 'W/m/m'}},
                        aux={'n_iter': 7, 'suppression': rng.standard_normal((3, 4)), 'cg': [(10, 1e-14)]})
    return OpenBTEResults(material=synthetic_material_2d(3, 4), mesh=geo, solvers={'bte': bte})


def test_state_round_trip(tmp_path):
    geo = pore_cell_2d(10.0, 0.25, 300, seed=1)
    r = _results(geo)
    fn = str(tmp_path / 'state')
    r.save(fn)
    r2 = OpenBTEResults.load(fn)
    assert r2.mesh.n_elems == geo.n_elems
    assert np.array_equal(r2.mesh.elem_volumes, geo.elem_volumes)
    assert set(r2.mesh.periodic_sides) == set(geo.periodic_sides)
    assert np.array_equal(r2.mesh.periodic_sides['Periodic_x'], geo.periodic_sides['Periodic_x'])
    assert r2.material.grid == r.material.grid and r2.material.h == r.material.h
    b1, b2 = r.solvers['bte'], r2.solvers['bte']
    assert b2.kappa_eff == b1.kappa_eff and b2.aux['n_iter'] == 7
    for k in b1.variables:
        assert np.array_equal(b1.variables[k]['data'], b2.variables[k]['data'])
        assert b1.variables[k]['units'] == b2.variables[k]['units']


def _read_vtk(fn):
    lines = open(fn).read().split('\n')
    i = lines.index([l for l in lines if l.startswith('POINTS')][0])
    n_pts = int(lines[i].split()[1])
    pts = np.array([list(map(float, l.split())) for l in lines[i + 1:i + 1 + n_pts]])
    j = lines.index([l for l in lines if l.startswith('CELLS')][0])
    n_cells = int(lines[j].split()[1])
    cells = np.array([list(map(int, l.split()))[1:] for l in lines[j + 1:j + 1 + n_cells]])
    data = {}
    for k, l in enumerate(lines):
        if l.startswith('SCALARS'):
            data[l.split()[1]] = np.array(lines[k + 2].split(), float)
        if l.startswith('VECTORS'):
            n = n_cells if 'CELL_DATA' in ''.join(lines) else n_pts
            data[l.split()[1]] = np.array([list(map(float, x.split())) for x in lines[k + 1:k + 1 + n]])
    return pts, cells, data


def test_vtk_cell_data_and_tiling(tmp_path):
    geo = pore_cell_2d(10.0, 0.25, 300, seed=1)
    r = _results(geo)
    fn = str(tmp_path / 'out')
    r.vtu(fn)
    pts, cells, data = _read_vtk(fn + '.vtk')
    assert pts.shape == (geo.n_nodes, 3) and np.allclose(pts[:, :2], geo.nodes) and np.all(pts[:, 2] == 0)
    assert np.array_equal(cells, geo.elems)
    assert np.allclose(data['Temperature_BTE[K]'], r.solvers['bte'].variables['Temperature_BTE']['data'])
    assert np.allclose(data['Flux_BTE[W/m/m]'][:, :2], r.solvers['bte'].variables['Flux_BTE']['data'])
    r.vtu(fn + '_2x1', repeat=[2, 1, 1])
    pts2, cells2, data2 = _read_vtk(fn + '_2x1.vtk')
    assert len(cells2) == 2 * geo.n_elems and len(pts2) == 2 * geo.n_nodes
    assert np.allclose(pts2[geo.n_nodes:, 0] - pts2[:geo.n_nodes, 0], geo.size[0])


def test_node_average_and_3d(tmp_path):
    geo = pore_slab_3d(10.0, 3.0, 0.2, 900, seed=2, nz=3)
    r = _results(geo)
    nodes, elems, cv = r._tiled([1, 1, 1])
    nv = OpenBTEResults.cell_to_node(cv, nodes, elems)
    # brute force on a few nodes
    T = cv['Temperature_BTE']['data']
    for n in [0, 17, 123]:
        owners = np.nonzero((elems == n).any(axis=1))[0]
        assert np.isclose(nv['Temperature_BTE']['data'][n], T[owners].mean())
    fn = str(tmp_path / 'o3')
    r.vtu(fn, dof='node', repeat=[1, 1, 2])
    pts, cells, data = _read_vtk(fn + '.vtk')
    assert len(cells) == 2 * geo.n_elems
    assert np.allclose(pts[geo.n_nodes:, 2] - pts[:geo.n_nodes, 2], geo.size[2])


def test_plot_over_line():
    geo = pore_cell_2d(10.0, 0.0, 400, seed=3)
    T = geo.elem_centroids[:, 0].copy()  # linear field
    res = OpenBTEResults(material=None, mesh=geo, solvers={'f': SolverResults(
        kappa_eff=0.0, variables={'T': {'data': T, 'units': 'K'}})})
    x = np.array([[-3.0, 0.0], [0.0, 0.0], [3.0, 0.0]])
    s, out = res.plot_over_line(['T'], x)
    assert np.allclose(s, [0, 3, 6])
    assert np.allclose(out['T'], x[:, 0], atol=0.3)
