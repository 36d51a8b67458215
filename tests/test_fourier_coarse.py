"""Host side of the two-level Fourier CG: the Morton renumbering gives the
same operator, gradient slots and contact cells in the new numbering, the
coarse inverse inverts the aggregated operator, and the two-level
preconditioner needs fewer CG iterations than Jacobi on the same system."""
import numpy as np
import scipy.sparse as sp

from openbte_amd.fourier import build_fourier_host, coarse_inverse, morton_order, permute_host
from openbte_amd.geometry import pore_cell_2d
from openbte_amd.material import synthetic_material_2d
from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity


def _host(n=6000):
    geo = pore_cell_2d(10.0 * (n / 2000) ** 0.5, 0.25, n, seed=3)
    mat = synthetic_material_2d(8, 8, 1.0, 1e4)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    return build_fourier_host(geo, np.asarray(mat.thermal_conductivity)[:2, :2], bcs, eff)


def _A(h):
    n = h.n
    r = np.repeat(np.arange(n), np.diff(h.row_ptr))
    return sp.csr_matrix((np.concatenate([h.diag, -h.w]), (np.concatenate([np.arange(n), r]),
                                                            np.concatenate([np.arange(n), h.col]))), shape=(n, n))


def _pcg(A, b, minv, tol=1e-12, maxit=5000):
    x = np.zeros_like(b)
    r = b.copy()
    z = minv(r)
    p = z.copy()
    rz = r @ z
    for it in range(maxit):
        if np.linalg.norm(r) <= tol * np.linalg.norm(b):
            return x, it
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        z = minv(r)
        rzn = r @ z
        p = z + rzn / rz * p
        rz = rzn
    return x, maxit


def test_morton_order_is_a_permutation_of_compact_aggregates():
    h = _host()
    perm = morton_order(h.xy)
    assert np.array_equal(np.sort(perm), np.arange(h.n))
    # aggregates of 256 consecutive cells are compact: their extent is far
    # below that of 256 cells taken in the original numbering
    xy = h.xy[perm]
    ext = lambda a: np.ptp(a[: (len(a) // 256) * 256].reshape(-1, 256, 2), axis=1).max(axis=1).mean()  # noqa: E731
    assert ext(xy) < 0.5 * ext(h.xy)


def test_permuted_system_is_the_same_system():
    h = _host()
    perm = morton_order(h.xy)
    hp = permute_host(h, perm)
    x = np.random.default_rng(1).standard_normal(h.n)
    np.testing.assert_allclose(_A(hp) @ x[perm], (_A(h) @ x)[perm], rtol=0, atol=1e-12 * np.abs(h.diag).max())
    assert perm[hp.pin] == h.pin
    np.testing.assert_array_equal(hp.B, h.B[perm])
    sc, scp = h.slot_cell.reshape(h.n, h.nv)[perm], hp.slot_cell.reshape(h.n, h.nv)
    assert np.all(np.where(sc >= 0, perm[np.maximum(scp, 0)] == sc, scp == -1))
    np.testing.assert_array_equal(hp.gweights.reshape(h.n, -1), h.gweights.reshape(h.n, -1)[perm])
    np.testing.assert_array_equal(perm[hp.side_e], h.side_e)
    np.testing.assert_array_equal(perm[hp.ce1], h.ce1)
    np.testing.assert_array_equal(perm[hp.ce2], h.ce2)


def test_coarse_inverse_and_two_level_iterations():
    h0 = _host(20000)
    h = permute_host(h0, morton_order(h0.xy))
    n, nc = h.n, (h.n + 255) // 256
    inv = coarse_inverse(h)
    Pm = sp.csr_matrix((np.ones(n), (np.arange(n), np.arange(n) >> 8)), shape=(n, nc))
    A = _A(h)
    Ac = (Pm.T @ A @ Pm).toarray()
    assert h.pin >= 0  # periodic cell: singular A, constant null space
    np.testing.assert_allclose((Ac + np.trace(Ac) / nc ** 2) @ inv, np.eye(nc), atol=1e-9)
    b = h.vol * h.B
    b = b - b.mean()  # consistent right-hand side of the singular system
    xj, itj = _pcg(A, b, lambda r: r / h.diag)
    x2, it2 = _pcg(A, b, lambda r: r / h.diag + Pm @ (inv @ (Pm.T @ r)))
    assert it2 < 0.75 * itj
    d = (x2 - x2.mean()) - (xj - xj.mean())
    assert np.abs(d).max() < 1e-9 * np.abs(xj - xj.mean()).max()
