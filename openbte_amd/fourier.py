"""``Fourier`` on MI355X: the diffusive initial guess (drop-in for
openbte/fourier.py:15-185).

The two-point-flux Laplacian with its orthogonal / non-orthogonal split
(openbte/objects.py:312-319) is assembled on the host into a symmetric CSR
operator (structure plus weights, once); every solve, gradient and
correction of the correction loop (openbte/fourier.py:149-178) runs as HIP
kernels (obte_fourier_*), with T, C and the gradient resident on the device.

Meshes of more than 3,072 cells (the one-workgroup solve takes the smaller
ones) are solved by CG preconditioned with one multigrid V-cycle per
iteration: the cells are renumbered along a Morton curve of their centroids,
every 8 consecutive cells form an aggregate, and a smoothed-aggregation
hierarchy (native obte_amg_create) goes down to at most 512 rows, whose dense
operator is inverted on the host.  Results are returned in the caller's cell
order.
"""
from __future__ import annotations

import ctypes as C
import warnings
from typing import NamedTuple, Optional

import numpy as np
import torch

from . import _lib
from .cache import FOURIER, content_key
from .objects import SolverResults

PINNED_CELL = 10  # openbte/fourier.py:43


def decomposed_directions(geo, kappa):
    """v_orth, v_nonorth of openbte/objects.py:312-319."""
    na = np.asarray(geo.normal_areas)
    v_orth = np.einsum('ki,ij,kj->k', na, kappa, na) / np.einsum('ki,ki->k', na, geo.dists)
    v_non = na @ kappa.T - geo.dists * v_orth[:, None]
    return v_orth, v_non


class FourierHost(NamedTuple):
    n: int
    dim: int
    nv: int
    pin: int
    row_ptr: np.ndarray
    col: np.ndarray
    w: np.ndarray
    diag: np.ndarray
    vol: np.ndarray
    B: np.ndarray
    slot_cell: np.ndarray
    slot_mode: np.ndarray
    slot_val: np.ndarray
    gweights: np.ndarray
    side_e: np.ndarray
    side_w: np.ndarray
    side_vn: np.ndarray
    slot_side: np.ndarray
    slot_sgn: np.ndarray
    contact_kind: int
    ce1: np.ndarray
    ce2: np.ndarray
    c_coef: np.ndarray
    c_value: float
    normalization: float
    xy: Optional[np.ndarray] = None   # [n][dim] cell centroids (the coarse level's ordering)


def build_fourier_host(geo, kappa, bcs, effective_thermal_conductivity=None, heat_source=(),
                       modified_fourier=False) -> FourierHost:
    """Symmetric operator A = V F and right-hand side B of openbte/fourier.py:46-131."""
    n, dim = int(geo.n_elems), int(geo.dim)
    esm = np.asarray(geo.elem_side_map)
    nv = esm.shape[1]
    sem = np.asarray(geo.side_elem_map)
    vol = np.asarray(geo.elem_volumes, np.float64)
    v_orth, v_non = decomposed_directions(geo, kappa)
    inn = np.asarray(geo.internal, np.int64)
    a, b = sem[inn, 0], sem[inn, 1]
    # per (cell, side slot): the neighbour across an internal side; a cell's
    # row of the operator is its internal slots in slot order (a kept
    # periodic side sits in the slots of both its cells)
    is_int = np.zeros(len(geo.sides), bool)
    is_int[inn] = True
    cells = np.repeat(np.arange(n), nv)
    sd = esm.reshape(-1)
    first = sem[sd, 0] == cells
    other = np.where(first, sem[sd, 1], sem[sd, 0])
    intr = is_int[sd]
    row_ptr = np.concatenate([[0], np.cumsum(intr.reshape(n, nv).sum(axis=1))])
    cols = other[intr]
    ws = v_orth[sd[intr]]
    diag = np.bincount(cells[intr], weights=ws, minlength=n)
    B = np.zeros(n)
    for key, contact in bcs.mixed.items():
        sides = np.asarray(geo.side_physical_regions[key], np.int64)
        ii = sem[sides, 0]
        h = contact['boundary_conductance']
        tmp = v_orth[sides] * h / (h + v_orth[sides] / geo.side_areas[sides])
        diag += np.bincount(ii, weights=tmp, minlength=n)
        B[ii] = contact['value'] * tmp / vol[ii]  # assignment (openbte/fourier.py:86)
    for region, jump in bcs.periodic.items():
        sides = np.asarray(geo.periodic_sides[region], np.int64)
        ii, jj = sem[sides, 0], sem[sides, 1]
        B += np.bincount(ii, weights=jump * v_orth[sides] / vol[ii], minlength=n)
        B -= np.bincount(jj, weights=jump * v_orth[sides] / vol[jj], minlength=n)
    if len(heat_source) != 0:
        B += (np.asarray(heat_source) * 1e-18) if not modified_fourier else np.asarray(heat_source)
    pin = -1
    if modified_fourier:
        diag = diag + vol
    elif len(bcs.mixed) == 0:
        pin = PINNED_CELL
        B[pin] = 0.0

    # gradient slots (openbte/objects.py:322-405)
    int_index = -np.ones(len(geo.sides), np.int64)
    int_index[inn] = np.arange(len(inn))
    side_jump = np.zeros(len(geo.sides))
    for region, sides in geo.periodic_sides.items():
        side_jump[np.asarray(sides, np.int64)] = float(bcs.periodic[region])
    mixed_val = np.full(len(geo.sides), np.nan)
    for key, value in bcs.mixed.items():
        mixed_val[np.asarray(geo.side_physical_regions[key], np.int64)] = float(value['value'])
    mode = np.zeros(n * nv, np.int32)
    val = np.zeros(n * nv)
    nb = -np.ones(n * nv, np.int64)
    intr = is_int[sd]
    mode[intr] = 1
    nb[intr] = other[intr]
    val[intr] = np.where(first[intr], side_jump[sd[intr]], -side_jump[sd[intr]])
    mix = ~intr & ~np.isnan(mixed_val[sd])
    mode[mix] = 2
    val[mix] = mixed_val[sd[mix]]
    slot_side = np.where(intr, int_index[sd], -1)
    slot_sgn = np.where(intr, np.where(first, 1.0, -1.0) / vol[cells], 0.0)

    contact_kind, ce1, ce2, c_coef, c_value, normalization = 0, np.zeros(0, np.int64), np.zeros(0, np.int64), \
        np.zeros(0), 0.0, 0.0
    eff = effective_thermal_conductivity
    if eff is not None:
        normalization = float(eff.normalization)
        if bcs.get_bc_by_name(eff.contact) == 'Mixed':
            sides = np.asarray(geo.side_physical_regions[eff.contact], np.int64)
            h = bcs.mixed[eff.contact]['boundary_conductance']
            contact_kind, ce1 = 2, sem[sides, 0]
            c_coef = v_orth[sides] * h / (h + v_orth[sides] / geo.side_areas[sides])
            c_value = float(bcs.mixed[eff.contact]['value'])
        else:
            sides = np.asarray(geo.periodic_sides[eff.contact], np.int64)
            contact_kind, ce1, ce2 = 1, sem[sides, 0], sem[sides, 1]
            c_coef = v_orth[sides]
            c_value = float(bcs.periodic[eff.contact])

    return FourierHost(
        n=n, dim=dim, nv=nv, pin=pin, row_ptr=row_ptr.astype(np.int32), col=cols.astype(np.int32), w=ws,
        diag=diag, vol=vol, B=B, slot_cell=nb.astype(np.int32), slot_mode=mode, slot_val=val,
        gweights=np.ascontiguousarray(np.asarray(geo.gradient_weights, np.float64)),
        side_e=np.ascontiguousarray(np.stack([a, b], 1).astype(np.int32)),
        side_w=np.asarray(geo.interp_weights, np.float64)[inn], side_vn=np.ascontiguousarray(v_non[inn]),
        slot_side=slot_side.astype(np.int32), slot_sgn=slot_sgn, contact_kind=contact_kind,
        ce1=np.asarray(ce1, np.int32), ce2=np.asarray(ce2, np.int32), c_coef=np.asarray(c_coef, np.float64),
        c_value=c_value, normalization=normalization,
        xy=np.ascontiguousarray(np.asarray(geo.elem_centroids, np.float64)[:, :dim]))


# Meshes above the one-workgroup solve (3,072 cells) get the multigrid
# preconditioner: cells renumbered along a Morton curve, aggregates of 8
# consecutive cells per level, levels down to at most 512 rows (dense inverse).
# Measured CG iterations of one solve from zero at 200k cells: 760 with the
# former two-level preconditioner (aggregates of 256 cells, Jacobi smoothing
# additive), 89 with the V-cycle.
MG_MIN_CELLS = 3073
MG_AGG = 8
MG_COARSE_MAX = 512
MG_OMEGA = 2.0 / 3.0   # prolongation smoothing weight (4/3 over the bound 2 of the spectrum of D^-1 A)
# prolongations are smoothed on levels with at most this many off-diagonal
# entries per row: 200k triangles smooth levels 0-1 (89 CG iterations from
# zero), 150k tetrahedra level 0 only (94; smoothing level 1 too gives 69 but
# a coarse level of ~520 entries per row)
MG_SMOOTH_MAX_ROW = 16
CG_CHECK_MG = 8


def _spread_bits(v: np.ndarray, dim: int) -> np.ndarray:
    """Interleave zeros between the bits of v (dim = 2: 16-bit inputs; dim =
    3: 10-bit inputs, two zeros between bits)."""
    v = v.astype(np.uint64)
    if dim == 2:
        for s, m in ((8, 0x00FF00FF), (4, 0x0F0F0F0F), (2, 0x33333333), (1, 0x55555555)):
            v = (v | (v << np.uint64(s))) & np.uint64(m)
    else:
        for s, m in ((16, 0x030000FF), (8, 0x0300F00F), (4, 0x030C30C3), (2, 0x09249249)):
            v = (v | (v << np.uint64(s))) & np.uint64(m)
    return v


def morton_order(xy: np.ndarray) -> np.ndarray:
    """Cell order along a Z (Morton) curve of the centroids (16 bits per axis
    in 2D, 10 in 3D; ties keep the original order)."""
    dim = xy.shape[1]
    bits = 16 if dim == 2 else 10
    q = ((xy - xy.min(0)) / (np.ptp(xy, 0) + 1e-300) * ((1 << bits) - 1)).astype(np.uint64)
    code = np.zeros(len(xy), np.uint64)
    for d in range(dim):
        code |= _spread_bits(q[:, d], dim) << np.uint64(d)
    return np.argsort(code, kind='stable')


def permute_host(h: FourierHost, perm: np.ndarray) -> FourierHost:
    """The same system with cell i of the result = cell perm[i] of ``h``."""
    n, nv = h.n, h.nv
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    counts = np.diff(h.row_ptr)[perm]
    row_ptr = np.concatenate([[0], np.cumsum(counts)])
    src = np.repeat(h.row_ptr[:-1][perm] - row_ptr[:-1], counts) + np.arange(row_ptr[-1])

    def cells(a):  # cell indices (-1 kept)
        a = np.asarray(a, np.int64)
        return np.where(a >= 0, inv[np.maximum(a, 0)], a).astype(np.int32)

    def rows(a, k):  # per-cell blocks of k values
        return np.ascontiguousarray(np.asarray(a).reshape(n, k)[perm].reshape(-1))
    return h._replace(
        pin=int(inv[h.pin]) if h.pin >= 0 else h.pin,
        row_ptr=row_ptr.astype(np.int32), col=cells(np.asarray(h.col)[src]), w=np.asarray(h.w)[src],
        diag=h.diag[perm], vol=h.vol[perm], B=h.B[perm],
        slot_cell=cells(rows(h.slot_cell, nv)), slot_mode=rows(h.slot_mode, nv), slot_val=rows(h.slot_val, nv),
        gweights=rows(h.gweights, h.dim * nv), side_e=cells(h.side_e), slot_side=rows(h.slot_side, nv),
        slot_sgn=rows(h.slot_sgn, nv), ce1=cells(h.ce1), ce2=cells(h.ce2),
        xy=None if h.xy is None else h.xy[perm])


class Hierarchy(NamedTuple):
    """The multigrid hierarchy of a (renumbered) Fourier system: per level
    above the coarsest its rows, operator (levels >= 1), prolongation P and
    restriction R = P^T as CSR arrays, and the inverse of the dense coarsest
    operator (plus a constant when the system is singular)."""
    levels: list
    coarse_inv: np.ndarray


def build_hierarchy(h: FourierHost) -> Hierarchy:
    """Smoothed-aggregation hierarchy of ``h`` (native ``obte_amg_create``)."""
    row_ptr = np.ascontiguousarray(h.row_ptr, np.int32)
    col = np.ascontiguousarray(h.col, np.int32)
    w = np.ascontiguousarray(h.w, np.float64)
    diag = np.ascontiguousarray(h.diag, np.float64)
    handle = C.c_void_p()
    _lib.call('obte_amg_create', h.n, _lib.ptr(row_ptr), _lib.ptr(col), _lib.ptr(w), _lib.ptr(diag), MG_AGG,
              MG_COARSE_MAX, MG_OMEGA, MG_SMOOTH_MAX_ROW, C.byref(handle))
    try:
        nl = C.c_int32(0)
        _lib.call('obte_amg_sizes', handle, C.byref(nl), None)
        sizes = np.zeros(4 * nl.value, np.int64)
        _lib.call('obte_amg_sizes', handle, C.byref(nl), _lib.ptr(sizes))
        sizes = sizes.reshape(-1, 4)
        levels = []
        for l in range(nl.value - 1):
            n, na, npv, nr = (int(v) for v in sizes[l])
            nn = int(sizes[l + 1, 0])
            lv = {'n': n,
                  'p_ptr': np.zeros(n + 1, np.int32), 'p_col': np.zeros(max(1, npv), np.int32),
                  'p_val': np.zeros(max(1, npv)), 'r_ptr': np.zeros(nn + 1, np.int32),
                  'r_col': np.zeros(max(1, nr), np.int32), 'r_val': np.zeros(max(1, nr))}
            a = [None] * 4
            if l > 0:
                lv.update({'a_ptr': np.zeros(n + 1, np.int32), 'a_col': np.zeros(max(1, na), np.int32),
                           'a_w': np.zeros(max(1, na)), 'a_diag': np.zeros(n)})
                a = [_lib.ptr(lv[k]) for k in ('a_ptr', 'a_col', 'a_w', 'a_diag')]
            _lib.call('obte_amg_export', handle, l, *a,
                      *[_lib.ptr(lv[k]) for k in ('p_ptr', 'p_col', 'p_val', 'r_ptr', 'r_col', 'r_val')])
            levels.append(lv)
        nc = int(sizes[-1, 0])
        dense = np.zeros((nc, nc))
        _lib.call('obte_amg_export_dense', handle, _lib.ptr(dense))
    finally:
        _lib.call('obte_amg_destroy', handle)
    if h.pin >= 0:
        # singular (no fixed temperature): a constant added to every entry
        # leaves the correction of zero-sum residuals unchanged
        dense += np.trace(dense) / (nc * nc)
    # a few BLAS threads only: preparation runs in several worker threads at
    # once (the serving loop), where all-core BLAS calls oversubscribe the host
    with _blas_threads(4):
        cinv = np.ascontiguousarray(np.linalg.inv(dense))
    return Hierarchy(levels=levels, coarse_inv=cinv)


def _blas_threads(k: int):
    try:
        from threadpoolctl import threadpool_limits
        return threadpool_limits(limits=k, user_api='blas')
    except ImportError:  # the limit is an optimisation only
        import contextlib
        return contextlib.nullcontext()


def prepare_host(host: FourierHost):
    """The host system as the device solves it: renumbered along a Morton
    curve with its multigrid hierarchy when the multigrid preconditioner
    applies (returns (host, perm or None, Hierarchy or None))."""
    if host.n >= MG_MIN_CELLS and host.xy is not None:
        perm = morton_order(host.xy)
        host = permute_host(host, perm)
        return host, perm, build_hierarchy(host)
    return host, None, None


def prepare(geometry, kappa, bcs, eff=None, heat_source=(), modified_fourier=False, global_temperature=0.0):
    """``prepare_host`` of the Fourier system of these inputs, cached by
    content (a solve on a mesh seen before skips the host work)."""
    g = geometry
    dict_arrays = []
    for dct in (g.periodic_sides, g.side_physical_regions):
        for k in sorted(dct):
            dict_arrays += [k, np.asarray(dct[k])]
    key = content_key('fourier', g.n_elems, g.dim, np.asarray(g.elem_side_map), np.asarray(g.side_elem_map),
                      np.asarray(g.normal_areas), np.asarray(g.elem_volumes), np.asarray(g.side_areas),
                      np.asarray(g.dists), np.asarray(g.gradient_weights), np.asarray(g.interp_weights),
                      np.asarray(g.elem_centroids), np.asarray(g.internal), *dict_arrays,
                      np.asarray(kappa, np.float64), repr(dict(bcs._asdict()) if hasattr(bcs, '_asdict') else bcs),
                      repr(eff), np.asarray(heat_source, np.float64), bool(modified_fourier),
                      float(global_temperature), MG_MIN_CELLS, MG_AGG, MG_COARSE_MAX, MG_OMEGA, MG_SMOOTH_MAX_ROW)

    def build():
        host = build_fourier_host(geometry, kappa, bcs, eff, heat_source, modified_fourier)
        if modified_fourier:
            host = host._replace(B=host.B + global_temperature)
        return prepare_host(host)
    return FOURIER.get_or_build(key, build)


class FourierDevice:
    """Device-resident Fourier operator and correction loop state."""

    def __init__(self, host: FourierHost, device=None, prepared=None):
        if not torch.cuda.is_available():
            raise RuntimeError('openbte_amd.Fourier needs a ROCm GPU; there is no CPU fallback')
        host, self.perm, hier = prepared if prepared is not None else prepare_host(host)
        self.h = host
        self.dev = dev = device or torch.device('cuda', torch.cuda.current_device())
        f64, i32 = torch.float64, torch.int32

        def T(a, dt):
            a = np.ascontiguousarray(a)
            if a.size == 0:
                a = np.zeros(1, dtype=a.dtype)
            return torch.as_tensor(a, dtype=dt).to(dev)
        h = host
        self.b = b = {}
        for name, dt in (('row_ptr', i32), ('col', i32), ('w', f64), ('diag', f64), ('vol', f64), ('B', f64),
                         ('slot_cell', i32), ('slot_mode', i32), ('slot_val', f64), ('gweights', f64),
                         ('side_e', i32), ('side_w', f64), ('side_vn', f64), ('slot_side', i32), ('slot_sgn', f64),
                         ('ce1', i32), ('ce2', i32), ('c_coef', f64)):
            b[name] = T(getattr(h, name), dt)
        n = h.n
        b['T'] = torch.zeros(n, dtype=f64, device=dev)
        b['C'] = torch.zeros(n, dtype=f64, device=dev)
        b['grad'] = torch.zeros(n * h.dim, dtype=f64, device=dev)
        b['ws'] = torch.zeros(6 * n + 64, dtype=f64, device=dev)
        b['scal'] = torch.zeros(16, dtype=f64, device=dev)
        b['side_tmp'] = torch.zeros(max(1, len(h.side_w)), dtype=f64, device=dev)
        b['red'] = torch.zeros(n // 256 + 2, dtype=f64, device=dev)
        b['out'] = torch.zeros(2, dtype=f64, device=dev)
        P = self.plan = _lib.FourierPlan()
        P.n, P.dim, P.nv, P.n_int, P.pin = n, h.dim, h.nv, len(h.side_w), h.pin
        for name in ('row_ptr', 'col', 'w', 'diag', 'vol', 'slot_cell', 'slot_mode', 'slot_val', 'gweights',
                     'side_e', 'side_w', 'side_vn', 'slot_side', 'slot_sgn', 'ce1', 'ce2', 'c_coef'):
            setattr(P, name, b[name].data_ptr())
        P.n_contact, P.contact_kind = len(h.ce1), h.contact_kind
        P.c_value, P.normalization = h.c_value, h.normalization
        self.levels_c = None
        if hier is not None:
            lv_c = (_lib.AmgLevel * len(hier.levels))()
            for l, lv in enumerate(hier.levels):
                L = lv_c[l]
                L.n = lv['n']
                L.a_entries = int(lv['a_ptr'][-1]) if 'a_ptr' in lv else int(h.row_ptr[-1])
                L.r_entries = int(lv['r_ptr'][-1])
                for k in ('a_ptr', 'a_col', 'p_ptr', 'p_col', 'r_ptr', 'r_col'):
                    if k in lv:
                        b[f'mg{l}_{k}'] = T(lv[k], i32)
                        setattr(L, k, b[f'mg{l}_{k}'].data_ptr())
                for k in ('a_w', 'a_diag', 'p_val', 'r_val'):
                    if k in lv:
                        b[f'mg{l}_{k}'] = T(lv[k], f64)
                        setattr(L, k, b[f'mg{l}_{k}'].data_ptr())
                for k in (('x', 'res') if l == 0 else ('b', 'x', 'res', 'z')):
                    b[f'mg{l}_{k}'] = torch.zeros(lv['n'], dtype=f64, device=dev)
                    setattr(L, k, b[f'mg{l}_{k}'].data_ptr())
            nc = hier.coarse_inv.shape[0]
            b['coarse_inv'] = T(hier.coarse_inv, f64)
            b['coarse_b'] = torch.zeros(nc, dtype=f64, device=dev)
            b['coarse_x'] = torch.zeros(nc, dtype=f64, device=dev)
            self.levels_c = lv_c  # the plan points at this host array
            P.n_levels, P.levels = len(hier.levels), C.addressof(lv_c)
            P.n_coarse, P.coarse_inv = nc, b['coarse_inv'].data_ptr()
            P.coarse_b, P.coarse_x = b['coarse_b'].data_ptr(), b['coarse_x'].data_ptr()
        self.stream = torch.cuda.current_stream(dev).cuda_stream
        # convergence test every CG_CHECK_MG iterations with the multigrid (an
        # iteration is ~12 launches, a test one host round trip), else every 16
        self.check_every = CG_CHECK_MG if hier is not None else 16
        self.cg_iters = []

    def in_cell_order(self, a: np.ndarray) -> np.ndarray:
        """Per-cell rows of the device numbering -> the caller's cell order."""
        if self.perm is None:
            return a
        out = np.empty_like(a)
        out[self.perm] = a
        return out

    def solve(self, rtol=1e-13, maxit=20000):
        b = self.b
        it = C.c_int32(0)
        rr = C.c_double(0.0)
        _lib.call('obte_fourier_solve', self.plan, b['B'].data_ptr(), b['C'].data_ptr(), b['T'].data_ptr(),
                  b['ws'].data_ptr(), b['scal'].data_ptr(), rtol, maxit, self.check_every, C.byref(it), C.byref(rr),
                  self.stream)
        self.cg_iters.append((it.value, rr.value))
        if rr.value > rtol:
            warnings.warn(f'Fourier CG stopped at {it.value} iterations with relative residual {rr.value:.2e} '
                          f'> {rtol:g}', RuntimeWarning, stacklevel=2)

    def gradient(self):
        _lib.call('obte_fourier_gradient', self.plan, self.b['T'].data_ptr(), self.b['grad'].data_ptr(), self.stream)

    def correction(self) -> float:
        b = self.b
        _lib.call('obte_fourier_correction', self.plan, b['grad'].data_ptr(), b['C'].data_ptr(),
                  b['side_tmp'].data_ptr(), b['red'].data_ptr(), b['out'].data_ptr(), self.stream)
        return float(np.sqrt(b['out'][0].item()))

    def kappa(self) -> float:
        _lib.call('obte_fourier_kappa', self.plan, self.b['T'].data_ptr(), self.b['out'][1:].data_ptr(),
                  self.stream)
        return float(self.b['out'][1].item())


def Fourier(geometry, thermal_conductivity, boundary_conditions, effective_thermal_conductivity=None,
            heat_source=[], global_temperature=0, modified_fourier=False, fourier_max_iter=40,
            fourier_rtol=1e-8, verbose=True, cg_rtol=1e-13) -> SolverResults:
    """Drop-in for ``openbte.fourier.Fourier`` (openbte/fourier.py:15-24)."""
    kappa = np.asarray(thermal_conductivity, np.float64)[:geometry.dim, :geometry.dim]
    prepared = prepare(geometry, kappa, boundary_conditions, effective_thermal_conductivity, heat_source,
                       modified_fourier, global_temperature)
    dev = FourierDevice(prepared[0], prepared=prepared)
    norm_C_old = 0.0
    error, n_iter, kappa_eff = 1.0, 0, 0.0
    while error > fourier_rtol and n_iter < fourier_max_iter:
        n_iter += 1
        dev.solve(rtol=cg_rtol)
        dev.gradient()
        norm_C = dev.correction()
        error = 0.0 if norm_C == 0 else abs(norm_C - norm_C_old) / norm_C
        norm_C_old = norm_C
        if effective_thermal_conductivity is not None:
            kappa_eff = dev.kappa()
        if verbose:
            print("Kappa (Fourier): {:E}, Iterations {:n}, Error {:E}".format(kappa_eff, n_iter, error))
    # the reference's ``while ... else`` resets kappa_eff (openbte/fourier.py:149,179-180)
    kappa_diffusive = kappa_eff
    T = dev.in_cell_order(dev.b['T'].cpu().numpy().copy())
    grad = dev.in_cell_order(dev.b['grad'].view(dev.h.n, dev.h.dim).cpu().numpy())
    J = -grad @ kappa.T
    variables = {'Temperature_Fourier': {'data': T, 'units': 'K'},
                 'Flux_Fourier': {'data': J * 1e9, 'units': 'W/m/m'}}
    return SolverResults(kappa_eff=0.0, variables=variables,
                         aux={'n_iter': n_iter, 'kappa_diffusive': kappa_diffusive, 'cg': dev.cg_iters,
                              'multigrid': dev.perm is not None})
