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
    The reference pairs two sides when they hold the two nodes of a
    ``$Periodic`` node pair and their centroids differ by that pair's offset
    (openbte/mesh.py:205-236).  Here the offsets of a ``$Periodic`` entity's
    node pairs give its translation (they must all agree: only translational
    periodicity is read), sides are paired by that translation, and every
    node pair must then join two sides so paired.  Without a ``$Periodic``
    section the translation is the difference of the two regions' mean side
    centroids.

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


def _periodic_entities(lines: List[str], dim: int):
    """Entities of a ``$Periodic`` section: (slave entity, master entity,
    node pairs [k][2] as (slave node id, master node id))."""
    out = []
    n_ent = int(lines[0].split()[0])
    i = 1
    for _ in range(n_ent):
        d, slave, master = (int(x) for x in lines[i].split()[:3])
        i += 1
        if lines[i].strip().startswith('Affine'):
            i += 1
        k = int(lines[i].split()[0])
        pairs = np.array(' '.join(lines[i + 1:i + 1 + k]).split(), dtype=np.int64).reshape(k, 2)
        i += 1 + k
        if d == dim - 1:
            out.append((slave, master, pairs))
    return out


def _periodic_from_section(lines, dim, nodes, index, face_nodes, face_tag, face_entity, region_names, tags):
    """Translation of every periodic region (``X_a`` onto ``X_b``) from the
    node pairs of the ``$Periodic`` section."""
    name_of_tag = dict(zip(tags, region_names))
    periodic: Dict[str, np.ndarray] = {}
    pairs_by_base: Dict[str, List[np.ndarray]] = {}
    for slave, master, pairs in _periodic_entities(lines, dim):
        ps = {name_of_tag[t] for t in face_tag[face_entity == slave].tolist()}
        if len(ps) != 1:
            raise ValueError(f'periodic entity {slave}: its faces carry {len(ps)} physical regions')
        sname = ps.pop()
        if not (sname.endswith('_a') or sname.endswith('_b')):
            continue
        off = nodes[index[pairs[:, 0]]] - nodes[index[pairs[:, 1]]]   # slave - master
        shift = off.mean(axis=0)
        if np.abs(off - shift).max() > 1e-9 * max(1.0, np.abs(nodes).max()):
            raise ValueError(f'periodic entity {slave}: node pairs are not one translation')
        base = sname[:-2]
        t = shift if sname.endswith('_b') else -shift               # X_a onto X_b
        if base in periodic and np.abs(periodic[base] - t).max() > 1e-9 * max(1.0, np.abs(nodes).max()):
            raise ValueError(f'periodic region {base}: entities disagree on the translation')
        periodic[base] = t
        pairs_by_base.setdefault(base, []).append(pairs)
    return periodic, pairs_by_base


def _check_node_pairs(mesh: Mesh, nodes, index, pairs_by_base) -> None:
    """Every node pair of the ``$Periodic`` section joins a kept side and its
    discarded partner: the pairing the reference derives from the same pairs."""
    for base, plist in pairs_by_base.items():
        kept = np.asarray(mesh.periodic_sides.get(base, []), np.int64)
        if len(kept) == 0:
            raise ValueError(f'periodic region {base}: no side pairs')
        kept_nodes = set(np.asarray(mesh.sides)[kept].reshape(-1).tolist())
        for pairs in plist:
            a, b = index[pairs[:, 0]], index[pairs[:, 1]]
            if not all((x in kept_nodes) or (y in kept_nodes) for x, y in zip(a.tolist(), b.tolist())):
                raise ValueError(f'periodic region {base}: a node pair joins no paired sides')


def read_msh(filename: str, periodic: Optional[Dict[str, np.ndarray]] = None) -> Mesh:
    """Read a gmsh 2.2 ASCII mesh.  ``periodic`` optionally gives the
    translation of each periodic region (``X_a`` onto ``X_b``); by default it
    comes from the ``$Periodic`` node pairs, or, without that section, from
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
    face_entity = np.array([r[4] if r[2] > 1 else -1 for r in face_rows], np.int64)

    # boundary side regions by node tuple
    tags = sorted(set(face_tag.tolist()))
    region_names = [names.get(t, (dim - 1, str(t)))[1] for t in tags]
    lookup = {tuple(fn): tags.index(t) for fn, t in zip(face_nodes.tolist(), face_tag.tolist())}

    def side_region(side_nodes, centroids):
        return np.array([lookup.get(tuple(sorted(s)), -1) for s in np.asarray(side_nodes).tolist()], np.int64)

    pairs_by_base = {}
    if periodic is None and 'Periodic' in sec:
        periodic, pairs_by_base = _periodic_from_section(sec['Periodic'], dim, nodes, index, face_nodes,
                                                         face_tag, face_entity, region_names, tags)
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
    mesh = build_mesh(nodes, elems, side_region, region_names, periodic=periodic, elem_regions=elem_regions)
    _check_node_pairs(mesh, nodes, index, pairs_by_base)
    return mesh


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
        # node pairs of every periodic region, slave = X_b (kept), master = X_a,
        # entity tags = the physical tags written above
        ent = []
        tag_of = {name: k + 1 for k, (d, name, _) in enumerate(regions)}
        for name, kept in mesh.periodic_sides.items():
            kept = np.asarray(kept)
            part = np.asarray(mesh.side_elem_map)[kept, 1]
            a_sides = [k for k, (d, nm, _) in enumerate(regions) if nm == name + '_a'][0]
            a_nodes = np.unique(np.asarray(mesh.sides)[regions[a_sides][2]].reshape(-1))
            b_nodes = np.unique(np.asarray(mesh.sides)[kept].reshape(-1))
            shift = mesh.side_centroids[kept].mean(axis=0) - \
                mesh.side_centroids[regions[a_sides][2]].mean(axis=0) if len(part) else np.zeros(dim)
            key = {tuple(np.round(mesh.nodes[i] + shift, 9).tolist()): i for i in a_nodes.tolist()}
            pairs = [(i, key[tuple(np.round(mesh.nodes[i], 9).tolist())]) for i in b_nodes.tolist()]
            ent.append((tag_of[name + '_b'], tag_of[name + '_a'], pairs))
        if ent:
            f.write(f'$Periodic\n{len(ent)}\n')
            for slave, master, pairs in ent:
                f.write(f'{dim - 1} {slave} {master}\n{len(pairs)}\n')
                for i, j in pairs:
                    f.write(f'{i + 1} {j + 1}\n')
            f.write('$EndPeriodic\n')
