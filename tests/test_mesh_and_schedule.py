"""Mesh generator contract (openbte/mesh.py:180-320 semantics) and the native
sweep scheduler."""
import numpy as np
import pytest

from openbte_amd.geometry import pore_cell_2d, pore_slab_3d, quad_cell_2d, structured_tri_cell_2d
from openbte_amd.material import synthetic_material_2d
from openbte_amd.objects import BoundaryConditions
from openbte_amd.operator import build_host_operator, build_schedule, fixed_width_inflow, level_pair_max, max_inflow


@pytest.mark.parametrize("geo", [pore_cell_2d(10.0, 0.25, 800, seed=1), quad_cell_2d(10.0, 6),
                                 structured_tri_cell_2d(10.0, 8),
                                 pore_cell_2d(20.0, 0.3, 1500, seed=2,
                                              pores=[('circle', (-4.0, -3.0), 3.0), ('square', (4.0, 4.0), 2.5)]),
                                 pore_slab_3d(10.0, 3.0, 0.2, 900, seed=2, nz=3)])
def test_mesh_contract(geo):
    sem = geo.side_elem_map
    na = geo.normal_areas
    # closed cells: sum of outward normal areas is zero
    S = np.zeros((geo.n_elems, geo.dim))
    act = np.concatenate([geo.internal, geo.boundary_sides])
    np.add.at(S, sem[act, 0], na[act])
    np.add.at(S, sem[geo.internal, 1], -na[geo.internal])
    assert np.abs(S).max() < 1e-12 * geo.side_areas.max()
    # normals point from side_elem_map[:,0] towards [:,1] on internal sides
    assert np.all(np.einsum('si,si->s', na[geo.internal], geo.dists[geo.internal]) > 0)
    # periodic pairs: kept side internal, partner element on the other edge
    L = geo.size[0]
    for name, sides in geo.periodic_sides.items():
        d = geo.dists[sides]
        axis = 0 if name.endswith('x') else 1
        assert np.all(np.abs(d[:, axis]) < 0.5 * L)
    # volumes add up
    assert geo.elem_volumes.min() > 0
    # interpolation weights in (0, 1)
    w = geo.interp_weights[geo.internal]
    assert w.min() > 0 and w.max() < 1


@pytest.mark.parametrize('max_width', [0, 7, 192])
def test_scheduler_topological_order(max_width):
    geo = pore_cell_2d(10.0, 0.25, 900, seed=3)
    mat = synthetic_material_2d(2, 24, 1.0, 1e4)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    op = build_host_operator(geo, mat, bcs)
    sc = build_schedule(op, max_width=max_width)
    free = build_schedule(op, max_width=0)
    widths = np.diff(sc.lvl_ptr)
    if max_width:
        # no level wider than the cap (padding entries between directions are negative)
        assert widths.max() <= max_width
    # capping the width only re-layers the cells: same cut faces per direction
    assert np.array_equal(np.diff(sc.cut_off), np.diff(free.cut_off))
    assert np.all(sc.n_levels >= free.n_levels)
    pair = level_pair_max(sc)
    lp0 = np.diff(sc.lvl_ptr[sc.lvl_off[0]:sc.lvl_off[0] + sc.n_levels[0] + 1])
    assert pair >= (lp0[:-1] + lp0[1:]).max()
    if max_width:
        assert pair <= 2 * max_width
    n = op.n
    kin = max(2, max_inflow(sc))
    src, face = fixed_width_inflow(sc, n, kin)
    # the scheduler's native slot export equals the host construction from the CSR
    assert sc.kin == kin and sc.nb_src is src
    src_h, face_h = fixed_width_inflow(sc._replace(nb_src=None, nb_face=None, kin=0), n, kin)
    assert np.array_equal(src, src_h) and np.array_equal(face, face_h)
    for l in range(op.n_dirs):
        lp = sc.lvl_ptr[sc.lvl_off[l]:sc.lvl_off[l + 1]]
        levelof = np.repeat(np.arange(len(lp) - 1), np.diff(lp))
        order = sc.order[l * n:(l + 1) * n]
        g = op.k_face @ op.dirs[l]
        used = face[l] >= 0
        # every inflow face (s.n < 0) appears once, periodic ones as interface rows
        assert used.sum() == (g < 0).sum()
        cons = np.broadcast_to(np.arange(n), src[l].shape)[used]
        s = src[l][used]
        f = face[l][used]
        assert np.all(op.i_mat[f] == order[cons])
        inner = s < n
        assert np.all(levelof[s[inner]] < levelof[cons[inner]])
        assert np.all(op.cut_candidate[f[~inner]] == 1)
        assert np.all(op.cut_candidate[f[inner]] == 0)
        ncut = sc.cut_off[l + 1] - sc.cut_off[l]
        assert s.max() < n + ncut


def test_scheduler_breaks_cycles():
    """A 3-cell ring with no cut candidates: the scheduler must cut a face
    and still produce a valid order."""
    from openbte_amd import operator as opm
    n = 3
    i_mat = np.array([0, 1, 2], np.int32)     # row cell
    j_mat = np.array([2, 0, 1], np.int32)     # upstream cell: 0<-2, 1<-0, 2<-1
    k = np.array([[-1.0, 0.0], [-1.0, 0.0], [-1.0, 0.0]])  # all inflow for s = (1, 0)
    host = opm.HostOperator(n=n, dim=2, n_mfp=1, n_dirs=1, i_mat=i_mat, j_mat=j_mat, k_face=k,
                            jump_face=np.zeros(3), cut_candidate=np.zeros(3, np.uint8),
                            cf_ptr=np.array([0, 1, 2, 3], np.int32), cf_face=np.arange(3, dtype=np.int32),
                            wall_ptr=np.zeros(4, np.int32), wall_side=np.zeros(0, np.int32),
                            wall_db=np.zeros((0, 2)), wall_temp=np.zeros(0), i_diffuse=np.zeros(0, np.int64),
                            wall_dbs=np.zeros((0, 2)), wall_rinv=np.zeros(0), contact_kind=0,
                            ce1=np.zeros(0, np.int32), ce2=np.zeros(0, np.int32), c_na=np.zeros((0, 2)),
                            contact_jump=0.0, normalization=0.0, dirs=np.array([[1.0, 0.0]]),
                            mfps=np.ones(1), t_coeff=np.ones((1, 1)), sigma=np.zeros((1, 1, 2)),
                            heat=np.zeros(3), nano=0.0)
    sc = build_schedule(host)
    assert sc.cut_off[-1] == 1
    assert sorted(sc.order.tolist()) == [0, 1, 2]
    assert sc.n_levels[0] == 3


def test_slab_3d_volume_and_regions():
    """Extruded tetrahedral slab: the tets tile the prism layers exactly, so
    the volume is (L^2 - pore polygon area) H; x and y sides pair up."""
    L, H, phi = 10.0, 3.0, 0.2
    geo = pore_slab_3d(L, H, phi, 900, seed=2, nz=3)
    base = pore_cell_2d(L, phi, int(round(900 / 9)), seed=2)
    assert geo.n_elems == 3 * 3 * base.n_elems
    assert abs(geo.elem_volumes.sum() - base.elem_volumes.sum() * H) < 1e-10 * L * L * H
    assert geo.dim == 3 and geo.elems.shape[1] == 4
    assert set(geo.periodic_sides) == {'Periodic_x', 'Periodic_y'}
    wall = geo.side_physical_regions['Boundary']
    z = geo.side_centroids[wall, 2]
    faces_z = np.isclose(z, 0.0) | np.isclose(z, H)
    # top and bottom faces: two triangles per base triangle each
    assert faces_z.sum() == 2 * base.n_elems
    assert np.allclose(geo.side_areas[wall][faces_z].sum(), 2 * base.elem_volumes.sum())
