"""CPU restatement of the mode interpolation onto (MFP, direction) grids
(TEST INFRASTRUCTURE; the oracle of ``openbte_amd.rta``).

Follows openbte/utils.py:281-337 (fast_interpolation, compute_polar,
compute_spherical) and openbte/material.py:7-102 (RTA3D), 105-189 (RTA2DSym)
with numpy and np.add.at, in the reference's order.

Pinning (tests/test_material_rta.py): interpolation weights sum to one per
mode, so the total weight and flux are conserved exactly; a mode on a grid
node lands on that node only; a mode between nodes splits in the closed-form
linear ratio; the database decode reproduces the stored bulk conductivity.
"""
from __future__ import annotations

import numpy as np


def fast_interpolation(fine, coarse, bound=False, scale='linear'):
    fine = np.array(fine, np.float64)
    coarse = np.asarray(coarse, np.float64)
    if scale == 'log':
        fine, coarse = np.log10(fine), np.log10(coarse)
    m2 = np.argmax(coarse >= fine[:, None], axis=1)
    m1 = m2 - 1
    a2 = (fine - coarse[m1]) / (coarse[m2] - coarse[m1])
    a1 = 1 - a2
    if bound == 'periodic':
        delta = coarse[-1] - coarse[-2]
        a = np.where(m2 == 0)[0]
        m1[a] = len(coarse) - 1
        m2[a] = 0
        fine[fine < delta / 2] += 2 * np.pi
        a2[a] = (fine[a] - coarse[-1]) / delta
        a1 = 1 - a2
    if bound == 'extent':
        al = np.where(fine < coarse[0])[0]
        m2[al] = 1
        m1[al] = 0
        a2[al] = (fine[al] - coarse[0]) / (coarse[1] - coarse[0])
        a1[al] = 1 - a2[al]
        ar = np.where(fine > coarse[-1])[0]
        m2[ar] = len(coarse) - 1
        m1[ar] = len(coarse) - 2
        a2[ar] = (fine[ar] - coarse[-2]) / (coarse[-1] - coarse[-2])
        a1[ar] = 1 - a2[ar]
    return a1, a2, m1 % len(coarse), m2


def _polar(mfp):
    phi = np.arctan2(mfp[:, 0], mfp[:, 1])
    phi[phi < 0] += 2 * np.pi
    return np.linalg.norm(mfp[:, :2], axis=1), phi


def rta2d_sym(C, tau, v, n_phi=48, n_mfp=50, mfp_min=None, mfp_max=None):
    """(W_sampled (M, L), sigma_sampled (M, L, 2), mfp grid (m), kappa, h, coeff)."""
    sigma = C[:, None] * v[:, :2]
    mfp = v[:, :2] * tau[:, None]
    f = np.divide(np.ones_like(tau), tau, out=np.zeros_like(tau), where=tau != 0)
    W = C * f
    r, phi_b = _polar(mfp)
    keep = np.where(r > 1e-10)
    r, mfp, phi_b, W, sigma = r[keep], mfp[keep], phi_b[keep], W[keep], sigma[keep]
    mfp_max = np.max(r) * 1.1 if mfp_max is None else mfp_max
    mfp_min = np.min(r) * 0.9 if mfp_min is None else mfp_min
    grid = np.logspace(np.log10(mfp_min) * 1.01, np.log10(mfp_max), n_mfp, endpoint=True)
    phi = np.linspace(0, 2.0 * np.pi, n_phi, endpoint=False)
    Ws = np.zeros((n_mfp, n_phi))
    Ss = np.zeros((n_mfp, n_phi, 2))
    a1, a2, m1, m2 = fast_interpolation(r, grid, bound='extent', scale='linear')
    b1, b2, p1, p2 = fast_interpolation(phi_b, phi, bound='periodic')
    for (mm, a) in ((m1, a1), (m2, a2)):
        for (pp, b) in ((p1, b1), (p2, b2)):
            np.add.at(Ws, (mm, pp), a * b * W)
            np.add.at(Ss, (mm, pp), sigma * (a * b)[:, None])
    inv = np.divide(1, Ws, out=np.zeros_like(Ws), where=Ws != 0)
    kappa = np.einsum('mqi,mq,mqj->ij', Ss, inv, Ss)
    h = 2 / 3 * np.sum(W * mfp[:, 0].clip(min=0)) * 1e-9
    return Ws, Ss, grid, kappa, h, 1 / np.sum(W)


def rta3d(C, tau, v, n_phi=24, n_mfp=50, n_theta=24):
    sigma = C[:, None] * v
    mfp = v * tau[:, None]
    f = np.divide(np.ones_like(tau), tau, out=np.zeros_like(tau), where=tau != 0)
    W = C * f
    r = np.linalg.norm(mfp, axis=1)
    keep = np.where(r > 1e-10)
    mfp, W, sigma, r = mfp[keep], W[keep], sigma[keep], r[keep]
    phi_b = np.arctan2(mfp[:, 0], mfp[:, 1])
    phi_b[phi_b < 0] += 2 * np.pi
    theta_b = np.arccos(mfp[:, 2] / r)
    dphi = 2.0 * np.pi / n_phi
    phi = np.linspace(dphi / 2.0, 2.0 * np.pi - dphi / 2.0, n_phi, endpoint=True)
    dth = np.pi / n_theta
    theta = np.linspace(dth / 2, np.pi - dth / 2, n_theta, endpoint=True)
    grid = np.logspace(np.log10(np.min(r) * 0.9) * 1.01, np.log10(np.max(r) * 1.1), n_mfp, endpoint=True)
    n_ang = n_phi * n_theta
    Ws = np.zeros((n_mfp, n_ang))
    Ss = np.zeros((n_mfp, n_ang, 3))
    a1, a2, m1, m2 = fast_interpolation(r, grid, bound='extent')
    b1, b2, p1, p2 = fast_interpolation(phi_b, phi, bound='periodic')
    c1, c2, t1, t2 = fast_interpolation(theta_b, theta, bound='extent')
    for (mm, a) in ((m1, a1), (m2, a2)):
        for (tt, c) in ((t1, c1), (t2, c2)):
            for (pp, b) in ((p1, b1), (p2, b2)):
                np.add.at(Ws, (mm, tt * n_phi + pp), a * b * c * W)
                np.add.at(Ss, (mm, tt * n_phi + pp), sigma * (a * b * c)[:, None])
    inv = np.divide(1, Ws, out=np.zeros_like(Ws), where=Ws != 0)
    kappa = np.einsum('mqi,mq,mqj->ij', Ss, inv, Ss)
    h = 2 / 3 * np.sum(W * mfp[:, 0].clip(min=0)) * 1e-18
    return Ws, Ss, grid, kappa, h, 1 / np.sum(W)
