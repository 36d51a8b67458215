"""gmsh ASCII 2.2 (.msh) meshes <-> the ``Mesh`` contract.

``read_msh`` replaces the reference's reader (``get_mesh``,
openbte/mesh.py:54-348) with array code: the reference walks every boundary
side through ``list.index`` (quadratic, openbte/mesh.py:142-144) and every
periodic node pair through Python sets; here sections are parsed in blocks
and all geometry comes from ``mesh.build_mesh`` (sides, normals out of
``side_elem_map[:,0]``, ``dists``, periodic pairing, gradient weights), so a
10^6-cell file reads in seconds.

Semantics kept from the reference:
  * the mesh is 2D when every node has z = 0 (openbte/mesh.py:73-79);
  * bulk elements are triangles / quadrangles (2D) or tetrahedra (3D), boundary
    elements are lines (2D) or triangles (3D) (openbte/mesh.py:89-92);
  * physical names label element regions and boundary side regions
    (openbte/mesh.py:127-145);
  * a periodic pair is the side regions ``<name>_a`` / ``<name>_b``; the
    ``_b`` side is kept and the ``_a`` side discarded (openbte/mesh.py:238-257).
    Sides are paired by the translation between the two regions, which is what
    the node pairs of the ``$Periodic`` section encode for translational
    periodicity; that section is therefore not needed and is skipped.

``write_msh`` writes the same format (nodes, physical names, bulk and
boundary elements with their physical tags) so that generated meshes can be
exchanged with gmsh-based tools.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

import numpy as np

from .mesh import build_mesh
from .objects import Mesh

# gmsh element types: 1 line, 2 triangle, 3 quadrangle, 4 tetrahedron
_BULK = {2: (2, 3), 3: (4,)}
_FACE = {2: (1,), 3: (2,)}


def _sections(text: str) -> Dict[str, List[str]]:
    out: Dict[str, List[str]] = {}
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        line = lines[i].strip()
        if line.startswith('$') and not line.startswith('$End'):
            name = line[1:]
            j = i + 1
            while j < len(lines) and lines[j].strip() != '$End' + name:
                j += 1
            out[name] = lines[i + 1:j]
            i = j
        i += 1
    return out


def read_msh(filename: str, periodic: Optional[Dict[str, np.ndarray]] = None) -> Mesh:
    """Read a gmsh 2.2 ASCII mesh.  ``periodic`` optionally gives the
    translation of each periodic region (``X_a`` onto ``X_b``); by default it is
    the difference of the two regions' mean side centroids."""
    if not filename.endswith('.msh'):
        filename = filename + '.msh'
    with open(filename) as f:
        sec = _sections(f.read())
    fmt = sec.get('MeshFormat', ['2.2 0 8'])[0].split()
    if not fmt[0].startswith('2'):
        raise ValueError(f'{filename}: gmsh format {fmt[0]} (only 2.x ASCII is read)')
    if len(fmt) > 1 and fmt[1] != '0':
        raise ValueError(f'{filename}: binary gmsh files are not read')
    names: Dict[int, Tuple[int, str]] = {}
    for line in sec.get('PhysicalNames', ['0'])[1:]:
        d, tag, name = line.split(maxsplit=2)
        names[int(tag)] = (int(d), name.strip().strip('"'))
    nl = sec['Nodes']
    n_nodes = int(nl[0])
    raw = np.array(' '.join(nl[1:1 + n_nodes]).split(), dtype=np.float64).reshape(n_nodes, 4)
    node_id = raw[:, 0].astype(np.int64)
    xyz = raw[:, 1:]
    index = np.full(node_id.max() + 1, -1, np.int64)
    index[node_id] = np.arange(n_nodes)
    dim = 2 if np.allclose(xyz[:, 2], 0.0) else 3
    nodes = np.ascontiguousarray(xyz[:, :dim])

    el = sec['Elements']
    n_el = int(el[0])
    rows = [np.array(l.split(), dtype=np.int64) for l in el[1:1 + n_el]]
    etype = np.array([r[1] for r in rows])
    bulk_rows = [r for r, t in zip(rows, etype) if t in _BULK[dim]]
    face_rows = [r for r, t in zip(rows, etype) if t in _FACE[dim]]
    if not bulk_rows:
        raise ValueError(f'{filename}: no {dim}D bulk elements')
    nv = {len(r) - 3 - r[2] for r in bulk_rows}
    if len(nv) != 1:
        raise ValueError(f'{filename}: mixed bulk element types are not supported')
    elems = index[np.array([r[3 + r[2]:] for r in bulk_rows])]
    bulk_tag = np.array([r[3] if r[2] > 0 else 0 for r in bulk_rows])
    face_nodes = (np.sort(index[np.array([r[3 + r[2]:] for r in face_rows])], axis=1)
                  if face_rows else np.zeros((0, dim), np.int64))
    face_tag = np.array([r[3] if r[2] > 0 else 0 for r in face_rows], np.int64)

    # boundary side regions by node tuple
    tags = sorted(set(face_tag.tolist()))
    region_names = [names.get(t, (dim - 1, str(t)))[1] for t in tags]
    lookup = {tuple(fn): tags.index(t) for fn, t in zip(face_nodes.tolist(), face_tag.tolist())}

    def side_region(side_nodes, centroids):
        return np.array([lookup.get(tuple(sorted(s)), -1) for s in np.asarray(side_nodes).tolist()], np.int64)

    # periodic translations from the _a / _b regions
    if periodic is None:
        periodic = {}
        for name in region_names:
            if name.endswith('_a') and name[:-2] + '_b' in region_names:
                base = name[:-2]
                ka, kb = tags[region_names.index(name)], tags[region_names.index(base + '_b')]
                ca = nodes[face_nodes[face_tag == ka]].mean(axis=(0, 1))
                cb = nodes[face_nodes[face_tag == kb]].mean(axis=(0, 1))
                periodic[base] = cb - ca
    elem_regions = {}
    for t in sorted(set(bulk_tag.tolist())):
        elem_regions[names.get(t, (dim, str(t)))[1]] = np.nonzero(bulk_tag == t)[0]
    return build_mesh(nodes, elems, side_region, region_names, periodic=periodic, elem_regions=elem_regions)


def get_mesh(filename: str = 'mesh') -> Mesh:
    """Name-compatible entry of the reference (openbte/mesh.py:54): reads
    ``filename + '.msh'``."""
    return read_msh(filename)


def write_msh(mesh: Mesh, filename: str) -> None:
    """Write ``mesh`` as gmsh 2.2 ASCII: nodes, physical names, bulk elements
    (physical region per ``elem_physical_regions``) and one boundary element
    per boundary side (physical region per ``side_physical_regions``; the two
    sides of every periodic pair as ``<name>_a`` / ``<name>_b``)."""
    if not filename.endswith('.msh'):
        filename = filename + '.msh'
    dim = mesh.dim
    nodes = np.zeros((mesh.n_nodes, 3))
    nodes[:, :dim] = mesh.nodes
    regions: List[Tuple[int, str, np.ndarray]] = []   # (dim, name, side or element indices)
    for name, sides in mesh.side_physical_regions.items():
        regions.append((dim - 1, name, np.asarray(sides)))
    # the discarded partner of a kept periodic side is neither internal nor a
    # boundary side; planar periodic regions are told apart by their normals
    n_sides = len(mesh.sides)
    disc = np.setdiff1d(np.arange(n_sides), np.concatenate([np.asarray(mesh.internal, np.int64),
                                                              np.asarray(mesh.boundary_sides, np.int64)]))
    unit = mesh.normal_areas / mesh.side_areas[:, None]
    free = np.ones(len(disc), bool)
    for name, kept in mesh.periodic_sides.items():
        kept = np.asarray(kept)
        nk = unit[kept[0]]
        sel = free & (np.abs(unit[disc] @ nk) > 1 - 1e-9)
        free &= ~sel
        regions.append((dim - 1, name + '_a', disc[sel]))
        regions.append((dim - 1, name + '_b', kept))
    for name, elems in mesh.elem_physical_regions.items():
        regions.append((dim, name, np.asarray(elems)))
    with open(filename, 'w') as f:
        f.write('$MeshFormat\n2.2 0 8\n$EndMeshFormat\n')
        f.write(f'$PhysicalNames\n{len(regions)}\n')
        for k, (d, name, _) in enumerate(regions):
            f.write(f'{d} {k + 1} "{name}"\n')
        f.write('$EndPhysicalNames\n')
        f.write(f'$Nodes\n{mesh.n_nodes}\n')
        for i, p in enumerate(nodes):
            f.write(f'{i + 1} {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n')
        f.write('$EndNodes\n')
        face_type = 1 if dim == 2 else 2
        nv = mesh.elems.shape[1]
        bulk_type = {3: 2, 4: 3}[nv] if dim == 2 else 4
        lines = []
        for k, (d, name, idx) in enumerate(regions):
            if d == dim - 1:
                for s in idx.tolist():
                    nds = ' '.join(str(int(x) + 1) for x in mesh.sides[s])
                    lines.append(f'{face_type} 2 {k + 1} {k + 1} {nds}')
        for k, (d, name, idx) in enumerate(regions):
            if d == dim:
                for e in idx.tolist():
                    nds = ' '.join(str(int(x) + 1) for x in mesh.elems[e])
                    lines.append(f'{bulk_type} 2 {k + 1} {k + 1} {nds}')
        f.write(f'$Elements\n{len(lines)}\n')
        for i, l in enumerate(lines):
            f.write(f'{i + 1} {l}\n')
        f.write('$EndElements\n')
