"""gmsh 2.2 reader/writer: round trips of generated meshes reproduce the
Mesh contract exactly; a hand-written file parses as gmsh lays it out."""
import os

import numpy as np
import pytest

from openbte_amd.geometry import pore_cell_2d, pore_slab_3d, quad_cell_2d
from openbte_amd.msh import get_mesh, read_msh, write_msh

FIELDS = ['n_elems', 'dim', 'sides', 'elems', 'side_elem_map', 'elem_side_map', 'normal_areas', 'elem_volumes',
          'side_areas', 'dists', 'internal', 'boundary_sides', 'gradient_weights', 'interp_weights',
          'second_order_correction']


@pytest.mark.parametrize("geo", [pore_cell_2d(10.0, 0.25, 800, seed=1), quad_cell_2d(10.0, 6),
                                 pore_slab_3d(10.0, 3.0, 0.2, 900, seed=2, nz=3)])
def test_round_trip(tmp_path, geo):
    fn = str(tmp_path / 'm.msh')
    write_msh(geo, fn)
    g2 = read_msh(fn)
    for f in FIELDS:
        a, b = np.asarray(getattr(geo, f)), np.asarray(getattr(g2, f))
        assert a.shape == b.shape, f
        assert np.allclose(a, b, rtol=1e-13, atol=1e-13), f
    for d1, d2 in [(geo.periodic_sides, g2.periodic_sides), (geo.side_physical_regions, g2.side_physical_regions)]:
        assert set(d1) == set(d2)
        for k in d1:
            assert np.array_equal(np.sort(d1[k]), np.sort(d2[k]))
    # the reference-style name without extension
    assert get_mesh(fn[:-4]).n_elems == geo.n_elems


def test_hand_written_file(tmp_path):
    """Unit square split into 2 triangles, periodic in x, walls in y; node ids
    not starting at 1 and an elementary tag after the physical one."""
    text = """$MeshFormat
2.2 0 8
$EndMeshFormat
$PhysicalNames
4
1 10 "Periodic_x_a"
1 11 "Periodic_x_b"
1 12 "Boundary"
2 20 "Bulk"
$EndPhysicalNames
$Nodes
4
11 0 0 0
12 1 0 0
13 1 1 0
14 0 1 0
$EndNodes
$Elements
6
1 1 2 10 1 14 11
2 1 2 11 2 12 13
3 1 2 12 3 11 12
4 1 2 12 4 13 14
5 2 2 20 5 11 12 13
6 2 2 20 5 11 13 14
$EndElements
"""
    fn = tmp_path / 'sq.msh'
    fn.write_text(text)
    g = read_msh(str(fn))
    assert g.n_elems == 2 and g.dim == 2
    assert np.allclose(g.elem_volumes, 0.5)
    assert len(g.periodic_sides['Periodic_x']) == 1
    assert len(g.side_physical_regions['Boundary']) == 2
    assert list(g.elem_physical_regions) == ['Bulk']
    # kept side is the x = 1 edge, its partner element is across the period
    s = g.periodic_sides['Periodic_x'][0]
    assert np.isclose(g.side_centroids[s][0], 1.0)
    assert g.side_elem_map[s][0] != g.side_elem_map[s][1]
