"""Results container, state files and VTK output.

``OpenBTEResults`` keeps the reference's container (openbte/objects.py:521-
541: material, mesh, solver results by name, optional mode data) and its
``get_variables``, ``vtu`` and ``plot_over_line`` (openbte/objects.py:545-
178 of the class).  State files differ on purpose: the reference pickles the
whole object into an SQLite dictionary (openbte/utils.py:243-257), which can
only be read back by executing the pickle; here ``save`` writes a NumPy
archive of plain arrays plus a JSON schema, and ``load`` reads it with
``allow_pickle=False``.

VTK output follows the reference's legacy ASCII unstructured grid (points,
cells of type 5 / 9 / 10, CELL_DATA or POINT_DATA with SCALARS and VECTORS
named ``key[units]``).  With ``repeat`` the periodic cell is tiled; the z
copies are shifted by the z size (the reference uses the y index for z,
openbte/objects.py:445, which stacks 3D copies on top of each other).
"""
from __future__ import annotations

import json
from typing import Dict, List, NamedTuple, Optional

import numpy as np

from .objects import Material, MaterialRTA, Mesh, SolverResults


def _flatten(prefix: str, obj, arrays: Dict[str, np.ndarray], schema):
    """Arrays into ``arrays`` under slash-joined keys; the structure (and the
    scalars) into a JSON-able schema."""
    if hasattr(obj, '_asdict'):
        return {'type': type(obj).__name__,
                'fields': {k: _flatten(f'{prefix}/{k}', v, arrays, schema) for k, v in obj._asdict().items()}}
    if isinstance(obj, dict):
        return {'type': 'dict', 'items': {str(k): _flatten(f'{prefix}/{k}', v, arrays, schema)
                                          for k, v in obj.items()}}
    if isinstance(obj, (list, tuple)) and not (len(obj) and isinstance(obj[0], (int, float, np.number))):
        return {'type': 'list', 'items': [_flatten(f'{prefix}/{i}', v, arrays, schema) for i, v in enumerate(obj)]}
    if obj is None or isinstance(obj, (str, bool, int, float)):
        return {'type': 'scalar', 'value': obj}
    if isinstance(obj, np.generic):
        return {'type': 'scalar', 'value': obj.item()}
    arrays[prefix] = np.asarray(obj)
    return {'type': 'array', 'key': prefix, 'tuple': isinstance(obj, tuple)}


_TYPES = {'Mesh': Mesh, 'Material': Material, 'MaterialRTA': MaterialRTA, 'SolverResults': SolverResults}


def _unflatten(node, arrays):
    t = node['type']
    if t == 'array':
        a = arrays[node['key']]
        return tuple(a.tolist()) if node.get('tuple') else a
    if t == 'scalar':
        return node['value']
    if t == 'dict':
        return {k: _unflatten(v, arrays) for k, v in node['items'].items()}
    if t == 'list':
        return [_unflatten(v, arrays) for v in node['items']]
    cls = _TYPES.get(t)
    fields = {k: _unflatten(v, arrays) for k, v in node['fields'].items()}
    if cls is None:
        raise ValueError(f'unknown record type {t}')
    return cls(**fields)


class OpenBTEResults(NamedTuple):
    material: Material
    mesh: Mesh
    solvers: dict = {}
    mat_mode: Optional[MaterialRTA] = None

    # -- state files -----------------------------------------------------------
    def save(self, filename: str = 'state') -> None:
        arrays: Dict[str, np.ndarray] = {}
        schema = _flatten('r', self, arrays, None)
        arrays['__schema__'] = np.frombuffer(json.dumps(schema).encode(), dtype=np.uint8)
        np.savez_compressed(filename if filename.endswith('.npz') else filename + '.npz', **arrays)

    @classmethod
    def load(cls, filename: str = 'state') -> 'OpenBTEResults':
        fn = filename if filename.endswith('.npz') else filename + '.npz'
        with np.load(fn, allow_pickle=False) as z:
            arrays = {k: z[k] for k in z.files}
        schema = json.loads(bytes(arrays.pop('__schema__')).decode())
        fields = {k: _unflatten(v, arrays) for k, v in schema['fields'].items()}
        return cls(**fields)

    # -- variables -------------------------------------------------------------
    def get_variables(self) -> dict:
        variables = {}
        for value in self.solvers.values():
            variables.update(value.variables)
        return variables

    def _tiled(self, repeat: List[int]):
        """Nodes, elements and copied cell variables of the tiled cell."""
        mesh = self.mesh
        variables = {k: {'data': np.asarray(v['data']), 'units': v['units']}
                     for k, v in self.get_variables().items()}
        repeat = list(repeat) + [1] * (3 - len(repeat))
        if repeat == [1, 1, 1]:
            return mesh.nodes, mesh.elems, variables
        shifts = []
        for x in range(repeat[0]):
            for y in range(repeat[1]):
                for z in range(repeat[2] if mesh.dim == 3 else 1):
                    shifts.append(np.array([x, y, z][:mesh.dim]) * np.asarray(mesh.size)[:mesh.dim])
        k = len(shifts)
        nodes = np.concatenate([mesh.nodes + s[None, :] for s in shifts])
        elems = np.concatenate([mesh.elems + i * mesh.n_nodes for i in range(k)])
        for v in variables.values():
            v['data'] = np.tile(v['data'], k) if v['data'].ndim == 1 else np.tile(v['data'], (k, 1))
        return nodes, elems, variables

    @staticmethod
    def cell_to_node(cell_variables: dict, nodes: np.ndarray, elems: np.ndarray) -> dict:
        """Node value = mean over the cells sharing the node (openbte/objects.py:461-489)."""
        n = len(nodes)
        conn = np.bincount(elems.reshape(-1), minlength=n).astype(np.float64)
        out = {}
        for key, v in cell_variables.items():
            d = np.asarray(v['data'])
            acc = np.zeros((n,) + d.shape[1:])
            np.add.at(acc, elems.reshape(-1), np.repeat(d, elems.shape[1], axis=0))
            out[key] = {'data': acc / (conn if d.ndim == 1 else conn[:, None]), 'units': v['units']}
        return out

    def vtu(self, filename: str = 'output', dof: str = 'cell', repeat: List[int] = [1, 1, 1]) -> None:
        """Legacy VTK file ``filename + '.vtk'`` with the solvers' variables."""
        mesh = self.mesh
        nodes, elems, variables = self._tiled(repeat)
        if dof != 'cell':
            variables = self.cell_to_node(variables, nodes, elems)
        n_nodes, n_elems, m = len(nodes), len(elems), elems.shape[1]
        ct = {2: {3: 5, 4: 9}, 3: {4: 10}}[mesh.dim][m]
        pts = np.zeros((n_nodes, 3))
        pts[:, :mesh.dim] = nodes
        out = ['# vtk DataFile Version 2.0', 'OpenBTE Data', 'ASCII', 'DATASET UNSTRUCTURED_GRID',
               f'POINTS {n_nodes} double']
        out += [' '.join(repr(float(x)) for x in p) for p in pts]
        out.append(f'CELLS {n_elems} {(m + 1) * n_elems}')
        out += [f'{m} ' + ' '.join(str(int(i)) for i in e) for e in elems]
        out.append(f'CELL_TYPES {n_elems}')
        out.append(' '.join([str(ct)] * n_elems))
        out.append(f'CELL_DATA {n_elems}' if dof == 'cell' else f'POINT_DATA {n_nodes}')
        for key, v in variables.items():
            name = f"{key}[{v['units']}]"
            d = np.asarray(v['data'])
            if d.ndim == 1:
                out += [f'SCALARS {name} double', 'LOOKUP_TABLE default', ' '.join(repr(float(x)) for x in d)]
            else:
                if d.shape[1] == 2:
                    d = np.concatenate([d, np.zeros((len(d), 1))], axis=1)
                out.append(f'VECTORS {name} double')
                out += [' '.join(repr(float(x)) for x in row) for row in d]
        with open(filename + '.vtk', 'w') as f:
            f.write('\n'.join(out) + '\n')

    def plot_over_line(self, variables: List[str], x: np.ndarray, repeat: List[int] = [1, 1, 1]):
        """Node-interpolated variables along the polyline ``x`` and its arc
        length (openbte/objects.py:556-583)."""
        from scipy.interpolate import griddata
        nodes, elems, cell_vars = self._tiled(repeat)
        node_vars = self.cell_to_node(cell_vars, nodes, elems)
        x = np.asarray(x, np.float64)
        out = {v: griddata(nodes, node_vars[v]['data'], x) for v in variables}
        s = np.concatenate([[0.0], np.cumsum(np.linalg.norm(np.diff(x, axis=0), axis=1))])
        return s, out
