"""Host side of the multigrid Fourier CG: the Morton renumbering gives the
same operator, gradient slots and contact cells in the new numbering; the
native hierarchy is the smoothed-aggregation construction (P = (I - w D^-1 A)
Pt, A_c = P^T A P) computed independently here with SciPy; the V-cycle the
device applies (restated below in NumPy) is a symmetric positive definite
preconditioner that needs far fewer CG iterations than Jacobi."""
import numpy as np
import scipy.sparse as sp

from openbte_amd import fourier as F
from openbte_amd.fourier import build_fourier_host, morton_order, permute_host
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


def _sa_reference(h, agg=F.MG_AGG, coarse_max=F.MG_COARSE_MAX, omega=F.MG_OMEGA, smooth_max=F.MG_SMOOTH_MAX_ROW):
    """Smoothed aggregation with SciPy: [(A, P)] per level and the coarsest A
    (P smoothed on levels with at most smooth_max off-diagonals per row)."""
    A = _A(h)
    out = []
    while A.shape[0] > coarse_max:
        n = A.shape[0]
        nc = (n + agg - 1) // agg
        Pt = sp.csr_matrix((np.ones(n), (np.arange(n), np.arange(n) // agg)), shape=(n, nc))
        off = A.nnz - np.count_nonzero(A.diagonal())
        P = (Pt - sp.diags(omega / A.diagonal()) @ (A @ Pt)).tocsr() if off <= smooth_max * n else Pt
        out.append((A, P))
        A = (P.T @ A @ P).tocsr()
    return out, A


def _level_A(lv, h):
    if 'a_ptr' not in lv:
        return _A(h)
    n = lv['n']
    r = np.repeat(np.arange(n), np.diff(lv['a_ptr']))
    return sp.csr_matrix((np.concatenate([lv['a_diag'], -lv['a_w'][:len(r)]]),
                          (np.concatenate([np.arange(n), r]), np.concatenate([np.arange(n), lv['a_col'][:len(r)]]))),
                         shape=(n, n))


def _csr(ptr, col, val, shape):
    m = ptr[-1]
    return sp.csr_matrix((val[:m], col[:m], ptr), shape=shape)


def _vcycle(levels, cinv, h, b, w=2.0 / 3.0):
    """The device V-cycle (k_mg_pre / k_mg_restrict / k_mg_coarse / k_mg_post)."""
    As = [_level_A(lv, h) for lv in levels]
    ns = [lv['n'] for lv in levels] + [cinv.shape[0]]
    Ps = [_csr(lv['p_ptr'], lv['p_col'], lv['p_val'], (ns[l], ns[l + 1])) for l, lv in enumerate(levels)]
    Rs = [_csr(lv['r_ptr'], lv['r_col'], lv['r_val'], (ns[l + 1], ns[l])) for l, lv in enumerate(levels)]

    def cyc(l, bl):
        if l == len(levels):
            return cinv @ bl
        d = As[l].diagonal()
        x = w * bl / d
        r = bl - As[l] @ x
        y = x + Ps[l] @ cyc(l + 1, Rs[l] @ r)
        return y + w * (bl - As[l] @ y) / d
    return cyc(0, b)


def test_hierarchy_is_smoothed_aggregation():
    h0 = _host(20000)
    h = permute_host(h0, morton_order(h0.xy))
    hier = F.build_hierarchy(h)
    ref, Ac = _sa_reference(h)
    assert len(hier.levels) == len(ref) >= 2
    for l, (lv, (A, P)) in enumerate(zip(hier.levels, ref)):
        nn = P.shape[1]
        Pn = _csr(lv['p_ptr'], lv['p_col'], lv['p_val'], P.shape)
        Rn = _csr(lv['r_ptr'], lv['r_col'], lv['r_val'], (nn, lv['n']))
        assert abs(Pn - P).max() <= 1e-12 * abs(P).max()
        assert abs(Rn - Pn.T).max() == 0.0
        An = _level_A(lv, h)
        assert abs(An - A).max() <= 1e-11 * abs(A).max()
    nc = Ac.shape[0]
    assert nc <= F.MG_COARSE_MAX
    assert h.pin >= 0  # periodic cell: singular operator, constant null space
    Ad = Ac.toarray()
    np.testing.assert_allclose((Ad + np.trace(Ad) / nc ** 2) @ hier.coarse_inv, np.eye(nc), atol=1e-8)


def test_vcycle_is_spd_and_cuts_cg_iterations():
    h0 = _host(20000)
    h = permute_host(h0, morton_order(h0.xy))
    hier = F.build_hierarchy(h)
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal((2, h.n))
    u -= u.mean()
    v -= v.mean()
    Mu, Mv = _vcycle(hier.levels, hier.coarse_inv, h, u), _vcycle(hier.levels, hier.coarse_inv, h, v)
    assert abs(u @ Mv - v @ Mu) <= 1e-10 * abs(u @ Mv)
    assert u @ Mu > 0 and v @ Mv > 0
    A = _A(h)
    b = h.vol * h.B
    b = b - b.mean()  # consistent right-hand side of the singular system
    xj, itj = _pcg(A, b, lambda r: r / h.diag)
    xm, itm = _pcg(A, b, lambda r: _vcycle(hier.levels, hier.coarse_inv, h, r))
    assert itm < 0.2 * itj, (itm, itj)
    d = (xm - xm.mean()) - (xj - xj.mean())
    assert np.abs(d).max() < 1e-9 * np.abs(xj - xj.mean()).max()


def test_hierarchy_3d_keeps_dense_levels_unsmoothed():
    """On tetrahedra the level-0 operator (~4 off-diagonals per row) gets a
    smoothed prolongation and the Galerkin level 1 (~40 per row) a piecewise
    constant one, so the coarse levels stay sparse; the hierarchy is the SciPy
    construction with the same rule."""
    from openbte_amd.geometry import pore_slab_3d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = pore_slab_3d(10.0, 4.0, 0.2, 20000, seed=4)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    h0 = build_fourier_host(geo, np.diag([150.0, 150.0, 150.0]), bcs, eff)
    h = permute_host(h0, morton_order(h0.xy))
    hier = F.build_hierarchy(h)
    ref, _ = _sa_reference(h)
    assert len(hier.levels) == len(ref) >= 2
    per_row = [np.diff(lv['p_ptr']).mean() for lv in hier.levels]
    assert per_row[0] > 2.0 and all(abs(p - 1.0) < 1e-12 for p in per_row[1:])
    for lv, (A, P) in zip(hier.levels, ref):
        Pn = _csr(lv['p_ptr'], lv['p_col'], lv['p_val'], P.shape)
        assert abs(Pn - P).max() <= 1e-12 * abs(P).max()
        assert abs(_level_A(lv, h) - A).max() <= 1e-11 * abs(A).max()
