"""Differentiation path: interfacial-transmission kappa and its adjoint
gradient on the GPU against the oracle (dense coupled solve, central
differences of it) and against central differences of the GPU forward."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(N, M, L, seed=0, mfp_max=1e2):
    from openbte_amd.geometry import structured_tri_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = structured_tri_cell_2d(10.0, N)
    mat = synthetic_material_2d(M, L, 1.0, mfp_max)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0})
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    pix = np.arange(geo.n_elems) // 4
    rho = np.random.default_rng(seed).uniform(0.2, 1.0, N * N)
    return geo, mat, bcs, eff, pix, rho


def test_tim_forward_and_gradient_match_oracle():
    from oracle import tim_oracle as to
    from openbte_amd.bte import SweepConfig
    from openbte_amd.tim import TIMSolver
    geo, mat, bcs, eff, pix, rho = _case(3, 2, 8)
    s = TIMSolver(geo, mat, bcs, eff, pix, cfg=SweepConfig(mc=4, block=64), rtol=1e-13)

    def dense(r):
        t, _, _, _, _, itf = to.side_transmission(geo, pix, r)
        return to.dense_kappa(to.assemble_tim(geo, mat, bcs, eff, t, itf))
    k = s.forward(rho)
    kd = dense(rho)
    print('kappa gpu', k, 'dense', kd, 'iters', s.n_iter)
    assert abs(k / kd - 1) < 1e-10
    g = s.gradient()
    h = 1e-6
    fd = np.array([(dense(rho + h * e) - dense(rho - h * e)) / (2 * h) for e in np.eye(len(rho))])
    print('grad', g, 'fd', fd)
    assert np.abs(g - fd).max() < 1e-6 * np.abs(fd).max()


def test_tim_plain_limit():
    """rho = 1 everywhere: the model is the plain solve."""
    from oracle import bte_oracle as bo
    from openbte_amd.bte import SweepConfig
    from openbte_amd.tim import TIMSolver
    geo, mat, bcs, eff, pix, _ = _case(3, 3, 8)
    s = TIMSolver(geo, mat, bcs, eff, pix, cfg=SweepConfig(mc=4, block=64), rtol=1e-13)
    k = s.forward(np.ones(9))
    r = bo.bte_rta(geo, mat, bcs, effective_thermal_conductivity=eff, bte_max_iter=2000, bte_rtol=1e-15)
    assert abs(k / r.kappa_eff - 1) < 1e-10


def test_tim_gradient_vs_gpu_fd_and_autograd():
    import torch
    from openbte_amd.bte import SweepConfig
    from openbte_amd.tim import KappaTIM, TIMSolver
    geo, mat, bcs, eff, pix, rho = _case(5, 6, 16, seed=3, mfp_max=1e3)
    s = TIMSolver(geo, mat, bcs, eff, pix, cfg=SweepConfig(mc=8, block=128), rtol=1e-13)
    r = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
    k = KappaTIM.apply(r, s)
    k.backward()
    g = r.grad.numpy()
    h = 1e-6
    for p in [0, 7, 12, 24]:
        e = np.zeros_like(rho)
        e[p] = h
        fd = (s.forward(rho + e) - s.forward(rho - e)) / (2 * h)
        print('pixel', p, 'adjoint', g[p], 'fd', fd)
        assert abs(g[p] - fd) < 1e-5 * np.abs(g).max()


def test_tim_full_mode_grid_gradient_and_recompute():
    """The differentiation workload's mode grid (30 MFP x 48 directions) on a
    6 x 6 pixel cell: kappa and the adjoint gradient against the oracle's
    exact reduced solves and their central differences
    (tests/golden/tim_6x6_modes.npz), to 1e-6 relative; the variant that
    recomputes the forward solutions a block of directions at a time instead
    of storing them gives the same gradient to 1e-12."""
    import os
    import sys
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
    sys.path.insert(0, here)
    from make_tim_modes import case
    from openbte_amd.tim import TIMSolver
    fx = np.load(os.path.join(here, 'tim_6x6_modes.npz'), allow_pickle=False)
    geo, mat, bcs, eff, pix, rho = case()
    assert np.array_equal(rho, fx['rho'])
    s = TIMSolver(geo, mat, bcs, eff, pix, rtol=1e-13)
    k = s.forward(rho)
    g = s.gradient()
    fd = fx['grad_fd']
    print('kappa', k, 'oracle', float(fx['kappa']), 'grad err', np.abs(g - fd).max() / np.abs(fd).max(), flush=True)
    assert abs(k / float(fx['kappa']) - 1) < 1e-10
    assert np.abs(g - fd).max() < 1e-6 * np.abs(fd).max()
    s2 = TIMSolver(geo, mat, bcs, eff, pix, rtol=1e-13, store_forward=False, dir_block=8)
    assert s2.forward(rho) == k
    assert s2.t_fwd is None
    g2 = s2.gradient()
    print('recomputed vs stored', np.abs(g2 - g).max() / np.abs(g).max(), flush=True)
    assert np.abs(g2 - g).max() <= 1e-12 * np.abs(g).max()
