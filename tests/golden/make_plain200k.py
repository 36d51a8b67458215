"""Writes tests/golden/plain200k_oracle.npz: the oracle's first outer
iterations of the 200k-cell workload (periodic square cell L = 100 nm,
centred circular pore of porosity 0.25, 199,713 seeded jittered-Delaunay
triangles, 60 MFP bins (1 nm - 100 um) x 96 directions, DeltaT = 1 K along
x), started from the oracle's own Fourier temperature.

Every one of the 5,760 (direction, MFP bin) systems is factorised by SuperLU
in every iteration (no LU cache: all factors would take ~200 GB), so one
iteration costs ~15,000 core-seconds; the systems of a direction run in one
forked worker and the partial sums are added in direction order.  The arrays
kept are the ones the GPU parity test compares (openbte/bte.py:222-249
semantics): DeltaT after each iteration, the per-mode kappa K and
suppression S, the wall temperatures TB, and the flux J on every 50th cell
plus its Frobenius norm.  Run from the repository root:

    python tests/golden/make_plain200k.py [workers] [iterations] [--resume]

--resume continues from the last iteration stored in the fixture (same mesh
hash): its DeltaT and TB are the state the next iteration starts from.
"""
import hashlib
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import bte_oracle as bo  # noqa: E402
from oracle import fourier_oracle as fo  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'plain200k_oracle.npz')
J_STRIDE = 50


def case():
    """The 200k-cell workload: same generator calls as the benchmark's default."""
    from openbte_amd.geometry import pore_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = pore_cell_2d(100.0, 0.25, 200_000, seed=0)
    mat = synthetic_material_2d(60, 96, 1.0, 1.0e5, seed=0)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    return geo, mat, bcs, eff


def mesh_hash(geo) -> str:
    h = hashlib.sha256()
    for a in (geo.nodes, geo.elems, geo.side_elem_map, geo.normal_areas, geo.elem_volumes):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


_OPS = None
_STATE = None


def _one_direction(l):
    """All MFP bins of direction l, factorised and solved (no cache)."""
    DeltaT, TB = _STATE
    w = bo._Worker(_OPS, np.array([l]), cache_lu=False)
    return l, tuple(w.iterate(DeltaT, TB))


def main():
    global _OPS, _STATE
    args = [a for a in sys.argv[1:] if a != '--resume']
    resume = '--resume' in sys.argv[1:]
    workers = int(args[0]) if len(args) > 0 else max(1, len(os.sched_getaffinity(0)) - 1)
    n_iter = int(args[1]) if len(args) > 1 else 2
    geo, mat, bcs, eff = case()
    t0 = time.perf_counter()
    f = fo.fourier(geo, mat.thermal_conductivity, bcs, eff)
    T_F = np.asarray(f.variables['Temperature_Fourier']['data'], np.float64)
    print(f'fourier {time.perf_counter() - t0:.1f}s, n={geo.n_elems}', flush=True)
    ops = bo.assemble(geo, mat, bcs, eff)
    _OPS = ops
    DeltaT = T_F.copy()
    TB = DeltaT[ops.i_diffuse].copy()
    out = {'mesh_sha256': np.array(mesh_hash(geo)), 'n_elems': geo.n_elems, 'n_mfp': ops.n_mfp,
           'n_dirs': ops.n_dirs, 'fourier_T_norm': np.linalg.norm(T_F), 'fourier_T_sum': T_F.sum(),
           'j_stride': J_STRIDE}
    first = 1
    if resume and os.path.exists(OUT):
        with np.load(OUT) as z:
            old = {k: z[k] for k in z.files}
        if str(old['mesh_sha256']) != str(out['mesh_sha256']):
            raise SystemExit('stored fixture is for another mesh')
        out = old
        first = int(old['n_iter']) + 1
        DeltaT = old[f'delta_t_{first - 1}'].copy()
        TB = old[f'tb_{first - 1}'].copy()
        print(f'resuming after iteration {first - 1}', flush=True)
    for it in range(first, n_iter + 1):
        t0 = time.perf_counter()
        _STATE = (DeltaT, TB)
        parts = [None] * ops.n_dirs
        ctx = mp.get_context('fork')
        with ctx.Pool(workers) as pool:
            for k, (l, s) in enumerate(pool.imap_unordered(_one_direction, range(ops.n_dirs))):
                parts[l] = bo.ModeSums(*s)
                if (k + 1) % 8 == 0:
                    print(f'  iteration {it}: {k + 1}/{ops.n_dirs} directions, '
                          f'{time.perf_counter() - t0:.0f}s', flush=True)
        DeltaT, TB, J, K, S = bo._reduce(parts)
        out[f'delta_t_{it}'] = DeltaT.copy()
        out[f'tb_{it}'] = TB.copy()
        out[f'K_{it}'] = K.copy()
        out[f'S_{it}'] = S.copy()
        out[f'kappa_{it}'] = K.sum()
        out[f'J_{it}_sample'] = J[::J_STRIDE].copy()
        out[f'J_{it}_norm'] = np.linalg.norm(J)
        out['n_iter'] = it
        np.savez_compressed(OUT, **out)
        print(f'iteration {it}: kappa {K.sum():.12f}, {time.perf_counter() - t0:.0f}s -> {OUT}', flush=True)


if __name__ == '__main__':
    main()
