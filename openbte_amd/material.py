"""Synthetic mode-resolved materials with the reference's ``Material`` semantics.

The reference interpolates ab-initio RTA data onto (MFP bin, direction) grids
(``RTA2DSym``, openbte/material.py:105-189).  The workloads here use a seeded
synthetic Si-like material with the same field meanings:

* directions ``s_l = [sin phi_l, cos phi_l]``, ``phi_l = 2 pi l / n_phi``
  (openbte/material.py:114-119), so l = 0 points along +y;
* ``t_coeff = gamma / sum(gamma)`` (openbte/material.py:167);
* ``sigma = C v * 1e-9`` and ``mfps`` in nm (openbte/material.py:183-189);
* ``thermal_conductivity = sum sigma sigma / gamma`` (openbte/material.py:164).
"""
from __future__ import annotations

import numpy as np

from .objects import Material


def _inv(gamma):
    """1/gamma with empty bins mapped to 0 (openbte/material.py:163)."""
    return np.divide(1.0, gamma, out=np.zeros_like(gamma), where=gamma != 0)


def kappa_weights(n_mfp: int, mfp_min: float, mfp_max: float, kappa_bulk: float = 150.0,
                  median: float = 300.0, sigma_ln: float = 1.2, seed: int = 0,
                  n_draws: int = 1_000_000):
    """MFP bins (nm, log-spaced) and bulk-kappa weights per bin.

    The weights are the histogram of ``n_draws`` lognormal MFP draws (median
    ``median`` nm, log-std ``sigma_ln``) over bins whose edges sit at the
    geometric mid-points of the MFP grid, with out-of-range draws clipped into
    the end bins, normalised so that they sum to ``kappa_bulk`` (W/m/K).
    """
    mfps = np.logspace(np.log10(mfp_min), np.log10(mfp_max), n_mfp)
    rng = np.random.default_rng(seed)
    draws = rng.lognormal(np.log(median), sigma_ln, n_draws)
    edges = np.sqrt(mfps[1:] * mfps[:-1]) if n_mfp > 1 else np.zeros(0)
    idx = np.searchsorted(edges, draws)
    w = np.bincount(idx, minlength=n_mfp).astype(np.float64)
    w *= kappa_bulk / w.sum()
    return mfps, w


def synthetic_material_2d(n_mfp: int = 30, n_phi: int = 48, mfp_min: float = 1.0,
                          mfp_max: float = 1.0e4, kappa_bulk: float = 150.0,
                          median: float = 300.0, sigma_ln: float = 1.2,
                          seed: int = 0) -> Material:
    """Si-like 2D material: ``n_mfp`` log-spaced MFP bins x ``n_phi`` directions.

    gamma_ml = 2 w_m / (Lambda_m^2 n_phi) (Lambda in m), so that
    kappa = sum_ml gamma Lambda^2 s s^T = kappa_bulk * I exactly.
    """
    mfps, w = kappa_weights(n_mfp, mfp_min, mfp_max, kappa_bulk, median, sigma_ln, seed)
    phi = 2 * np.pi * np.arange(n_phi) / n_phi
    s = np.stack([np.sin(phi), np.cos(phi)], axis=1)
    lam = mfps * 1e-9
    gamma = (2.0 * w / lam ** 2)[:, None] / n_phi * np.ones((1, n_phi))
    sigma = gamma[:, :, None] * lam[:, None, None] * s[None, :, :]
    kappa = np.einsum('mli,ml,mlj->ij', sigma, _inv(gamma), sigma)
    t_coeff = gamma / gamma.sum()
    h = 2.0 / 3.0 * np.sum(gamma * (lam[:, None] * s[None, :, 0]).clip(min=0)) * 1e-9
    return Material(thermal_conductivity=kappa,
                    sigma=sigma * 1e-9,
                    vmfp=s,
                    t_coeff=t_coeff,
                    mfps=mfps.copy(),
                    grid=(n_mfp, n_phi),
                    h=h,
                    heat_source_coeff=1.0 / gamma.sum())


def synthetic_material_3d(n_mfp: int = 20, n_theta: int = 8, n_phi: int = 8,
                          mfp_min: float = 1.0, mfp_max: float = 1.0e4,
                          kappa_bulk: float = 150.0, median: float = 300.0,
                          sigma_ln: float = 1.2, seed: int = 0) -> Material:
    """Si-like 3D material on an n_theta x n_phi direction grid.

    Polar cells are uniform in cos(theta) so every direction has solid angle
    weight 1/(n_theta n_phi); direction l = t * n_phi + p (theta-major, as the
    3D grid of openbte/material.py:10-28).  gamma_ml = 3 w_m omega_l / Lambda_m^2
    gives kappa = kappa_bulk * <3 s s^T>, isotropic up to the quadrature.
    """
    mfps, w = kappa_weights(n_mfp, mfp_min, mfp_max, kappa_bulk, median, sigma_ln, seed)
    mu = -1 + (2 * np.arange(n_theta) + 1) / n_theta
    th = np.arccos(mu)
    ph = 2 * np.pi * (np.arange(n_phi) + 0.5) / n_phi
    T, P = np.meshgrid(th, ph, indexing='ij')
    s = np.stack([np.sin(T) * np.sin(P), np.sin(T) * np.cos(P), np.cos(T)], -1).reshape(-1, 3)
    n_dirs = n_theta * n_phi
    lam = mfps * 1e-9
    gamma = (3.0 * w / lam ** 2)[:, None] / n_dirs * np.ones((1, n_dirs))
    sigma = gamma[:, :, None] * lam[:, None, None] * s[None, :, :]
    kappa = np.einsum('mli,ml,mlj->ij', sigma, _inv(gamma), sigma)
    t_coeff = gamma / gamma.sum()
    return Material(thermal_conductivity=kappa,
                    sigma=sigma * 1e-9,
                    vmfp=s,
                    t_coeff=t_coeff,
                    mfps=mfps.copy(),
                    grid=(n_mfp, n_theta, n_phi),
                    h=0.0,
                    heat_source_coeff=1.0 / gamma.sum())


def gray_2d(mfp: float, n_phi: int = 96, kappa: float = 1.0) -> Material:
    """Single-MFP gray material with the semantics of ``Gray2D``
    (openbte/material.py:217-247): sigma = 2 kappa / MFP * s / n_phi,
    t = 1/n_phi.  Directions are not rounded."""
    phi = 2 * np.pi * np.arange(n_phi) / n_phi
    s = np.stack([np.sin(phi), np.cos(phi)], axis=1)
    sigma = np.zeros((1, n_phi, 2))
    sigma[0] = 2 * kappa / mfp * s / n_phi
    return Material(thermal_conductivity=kappa * np.eye(2),
                    sigma=sigma,
                    vmfp=s,
                    t_coeff=np.ones((1, n_phi)) / n_phi,
                    mfps=np.array([float(mfp)]),
                    grid=(1, n_phi),
                    h=4 / np.pi * kappa / mfp,
                    heat_source_coeff=mfp * mfp * 0.5 / kappa)
