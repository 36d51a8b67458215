"""Writes tests/golden/tim_6x6_modes.npz: kappa_eff of the interfacial-
transmission model and its central-difference gradient with respect to the
36 pixel densities of a 6 x 6 pixel cell (144 triangles, the pixel size of
the 40 x 40 differentiation workload: 2.5 nm) at that workload's full mode
grid, 30 MFP bins (1 nm - 10 um) x 48 directions.  Every kappa is the
oracle's exact reduced fixed-point solve (oracle/tim_oracle.py:
reduced_kappa, no iteration error), so the differences (h = 1e-5) are
accurate to ~1e-9 relative.  Run from the repository root:

    OPENBLAS_NUM_THREADS=1 python tests/golden/make_tim_modes.py [workers]
"""
import multiprocessing as mp
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import tim_oracle as to  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tim_6x6_modes.npz')
N_PIX = 6
H = 1e-5


def case():
    """6 x 6 pixels of 2.5 nm, 4 triangles per pixel, 30 MFP x 48 directions,
    periodic, kappa along x; seeded densities in [0.2, 1]."""
    from openbte_amd.geometry import structured_tri_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = structured_tri_cell_2d(2.5 * N_PIX, N_PIX)
    mat = synthetic_material_2d(30, 48, 1.0, 1.0e4, seed=11)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0})
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    pix = np.arange(geo.n_elems) // 4
    rho = np.random.default_rng(17).uniform(0.2, 1.0, N_PIX * N_PIX)
    return geo, mat, bcs, eff, pix, rho


def kappa(r):
    geo, mat, bcs, eff, pix, _ = case()
    t, _, _, _, _, itf = to.side_transmission(geo, pix, r)
    return to.reduced_kappa(to.assemble_tim(geo, mat, bcs, eff, t, itf))


def main():
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 1
    _, _, _, _, _, rho = case()
    pts = [rho]
    for p in range(len(rho)):
        e = np.zeros_like(rho)
        e[p] = H
        pts += [rho + e, rho - e]
    with mp.get_context('fork').Pool(workers) as pool:
        ks = np.array(pool.map(kappa, pts))
    fd = (ks[1::2] - ks[2::2]) / (2 * H)
    np.savez_compressed(OUT, rho=rho, kappa=ks[0], grad_fd=fd, h=H)
    print(f'{OUT}: kappa {ks[0]:.12f}, |grad| max {np.abs(fd).max():.4e}')


if __name__ == '__main__':
    main()
