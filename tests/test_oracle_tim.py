"""Pinning of the interfacial-transmission oracle (CPU): the Jacobi outer
iteration equals the dense coupled solve, t = 1 is the plain model, and the
model conserves energy across interfaces (the contact flux equals the flux
through any other full cross-section of faces)."""
import numpy as np

from oracle import bte_oracle as bo, tim_oracle as to


def _case(N=3, M=2, L=8, seed=0):
    from openbte_amd.geometry import structured_tri_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = structured_tri_cell_2d(10.0, N)
    mat = synthetic_material_2d(M, L, 1.0, 1e2)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0})
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    pix = np.arange(geo.n_elems) // 4
    rho = np.random.default_rng(seed).uniform(0.2, 1.0, N * N)
    return geo, mat, bcs, eff, pix, rho


def test_iteration_equals_dense_fixed_point():
    geo, mat, bcs, eff, pix, rho = _case()
    t, _, _, _, _, itf = to.side_transmission(geo, pix, rho)
    T = to.assemble_tim(geo, mat, bcs, eff, t, itf)
    kd = to.dense_kappa(T)
    k, _, _, _ = to.iterate(T, np.zeros(geo.n_elems), 1000, 1e-15)
    assert abs(k / kd - 1) < 1e-11


def test_full_transmission_is_plain_model():
    geo, mat, bcs, eff, pix, rho = _case(M=3)
    t, _, _, _, _, itf = to.side_transmission(geo, pix, np.ones_like(rho))
    assert np.all(t == 1.0)
    k1 = to.dense_kappa(to.assemble_tim(geo, mat, bcs, eff, t, itf))
    r = bo.bte_rta(geo, mat, bcs, effective_thermal_conductivity=eff, bte_max_iter=2000, bte_rtol=1e-15)
    assert abs(k1 / r.kappa_eff - 1) < 1e-11


def test_interpolation_and_derivatives():
    """t = harmonic mean of the regularised densities; derivatives by central differences."""
    geo, mat, bcs, eff, pix, rho = _case(N=4)
    t, da, db, pa, pb, itf = to.side_transmission(geo, pix, rho)
    r = to.K0 + rho * (1 - to.K0)
    assert np.allclose(t[itf], 2 / (1 / r[pa[itf]] + 1 / r[pb[itf]]), rtol=1e-14)
    assert np.all(t[~itf] == 1.0)
    h = 1e-7
    for p in range(len(rho)):
        e = np.zeros_like(rho)
        e[p] = h
        tp = to.side_transmission(geo, pix, rho + e)[0]
        tm = to.side_transmission(geo, pix, rho - e)[0]
        fd = (tp - tm) / (2 * h)
        an = da * (pa == p) + db * (pb == p)
        assert np.abs(fd - an).max() < 1e-7


def test_energy_conservation_across_interfaces():
    """Steady state without sources: the net x-flux through the contact equals
    the net x-flux through the column of faces one pixel further in, where
    the flux through a face is t times the upwind transmitted part (the
    reflected part does not cross)."""
    geo, mat, bcs, eff, pix, rho = _case(N=3, M=2, L=8, seed=5)
    t, _, _, _, _, itf = to.side_transmission(geo, pix, rho)
    T = to.assemble_tim(geo, mat, bcs, eff, t, itf)
    ops = T.ops
    n = geo.n_elems
    # converged mode solutions
    k, DeltaT, TB, _ = to.iterate(T, np.zeros(n), 2000, 1e-15)
    inn = np.asarray(geo.internal)
    sem = geo.side_elem_map
    xs = geo.side_centroids[:, 0]
    side_t = np.ones(len(geo.sides))
    side_t[inn] = t

    def plane_flux(sides, jump):
        tot = 0.0
        for l in range(ops.n_dirs):
            for m in range(ops.n_mfp):
                A = bo.system_matrix(ops, l, m)
                x = np.linalg.solve(A.toarray(), to._rhs(T, l, m, DeltaT, TB))
                sn = np.einsum('j,cj->c', mat.sigma[m, l, :2], geo.normal_areas[sides])
                up = np.where(sn > 0, x[sem[sides, 0]], x[sem[sides, 1]] + jump)
                tot += np.sum(side_t[sides] * sn * up)
        return tot
    contact = np.asarray(geo.periodic_sides['Periodic_x'])
    f_contact = plane_flux(contact, 1.0)
    L = geo.size[0]
    h = L / 3
    inner = inn[np.isclose(xs[inn], -L / 2 + h)]
    # orient normals along +x
    f_inner = plane_flux(inner, 0.0) * np.sign(geo.normal_areas[inner[0], 0])
    assert abs(f_inner / f_contact - 1) < 1e-9


def test_reduced_solve_equals_dense_solve():
    """The reduced (DeltaT, TB) fixed-point solve equals the dense coupled
    solve and the converged Jacobi iteration (so it can stand in for them at
    mode counts the dense solve cannot reach)."""
    geo, mat, bcs, eff, pix, rho = _case(N=3, M=2, L=8, seed=5)
    t, _, _, _, _, itf = to.side_transmission(geo, pix, rho)
    T = to.assemble_tim(geo, mat, bcs, eff, t, itf)
    kd = to.dense_kappa(T)
    kr = to.reduced_kappa(T)
    assert abs(kr / kd - 1) < 1e-11
    k, _, _, _ = to.iterate(T, np.zeros(geo.n_elems), 2000, 1e-15)
    assert abs(kr / k - 1) < 1e-10
